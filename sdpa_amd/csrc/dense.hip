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
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

namespace sdpa_amd {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TB = 64;    // tile edge
constexpr int TLD = 65;   // LDS leading dimension (pad)
constexpr int BK = 16;    // k-depth per LDS stage in the streaming GEMM

// Work-group barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global load, which would stall the compute waves on the
// tile they have just started to prefetch; nothing passes between the waves
// of these kernels through global memory (the tagged words carry their own
// validity).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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
  // this thread's elements of the A tile, op(A)[bm + r][k0 + k] (r < 64,
  // k < 16) -> As[r + 64 k], and of the B tile, op(B)[k0 + k][bn + c] ->
  // Bs[k + 16 c]; the next step's are loaded before the current step's
  // product, so the loads overlap it
  auto load_tiles = [&](int k0, double (&pa)[4], double (&pb)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;  // 0..1023
      int r, k;
      if (TA == 0) { r = idx & 63; k = idx >> 6; }
      else { k = idx & 15; r = idx >> 4; }
      const int gr = bm + r, gk = k0 + k;
      pa[e] = (gr < it.M && gk < it.K)
                  ? (TA == 0 ? it.A[gr + (size_t)it.lda * gk] : it.A[gk + (size_t)it.lda * gr])
                  : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      int k, c;
      if (TBt == 0) { k = idx & 15; c = idx >> 4; }
      else { c = idx & 63; k = idx >> 6; }
      const int gk = k0 + k, gc = bn + c;
      pb[e] = (gk < it.K && gc < it.N)
                  ? (TBt == 0 ? it.B[gk + (size_t)it.ldb * gc] : it.B[gc + (size_t)it.ldb * gk])
                  : 0.0;
    }
  };
  double pa[4], pb[4];
  load_tiles(0, pa, pb);
  for (int k0 = 0; k0 < it.K; k0 += BK) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      if (TA == 0) As[(idx & 63) + 64 * (idx >> 6)] = pa[e];
      else As[(idx >> 4) + 64 * (idx & 15)] = pa[e];
      if (TBt == 0) Bs[(idx & 15) + 16 * (idx >> 4)] = pb[e];
      else Bs[(idx >> 6) + 16 * (idx & 63)] = pb[e];
    }
    __syncthreads();
    if (k0 + BK < it.K) load_tiles(k0 + BK, pa, pb);
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

constexpr size_t kDescPoolMax = 64;  // copies kept before the pool starts over

int DenseWork::desc_copy(hipStream_t s, const void* h, size_t bytes, const void** out) {
  for (const DescCopy& e : desc_pool)
    if (e.bytes.size() == bytes && std::memcmp(e.bytes.data(), h, bytes) == 0) {
      *out = e.d;
      return 0;
    }
  if (desc_pool.size() >= kDescPoolMax) {
    // launches on either stream may still read the old copies
    SDPA_HIP_CHECK(hipStreamSynchronize(s));
    if (side) SDPA_HIP_CHECK(hipStreamSynchronize(side));
    for (DescCopy& e : desc_pool) SDPA_HIP_CHECK(hipFree(e.d));
    desc_pool.clear();
  }
  DescCopy e;
  e.bytes.assign(static_cast<const unsigned char*>(h), static_cast<const unsigned char*>(h) + bytes);
  SDPA_HIP_CHECK(hipMalloc(&e.d, bytes));
  SDPA_HIP_CHECK(hipMemcpyAsync(e.d, h, bytes, hipMemcpyHostToDevice, s));
  *out = e.d;
  desc_pool.push_back(std::move(e));
  return 0;
}

int DenseWork::ensure_dinv(size_t n) {
  if (n <= cap_dinv) return 0;
  if (d_dinv) SDPA_HIP_CHECK(hipFree(d_dinv));
  cap_dinv = n;
  SDPA_HIP_CHECK(hipMalloc(&d_dinv, sizeof(double) * cap_dinv));
  return 0;
}
int DenseWork::ensure_flags(int n) {
  if (!d_err) {
    SDPA_HIP_CHECK(hipMalloc(&d_err, sizeof(int)));
    SDPA_HIP_CHECK(hipMemset(d_err, 0, sizeof(int)));
    SDPA_HIP_CHECK(hipHostMalloc(&h_err, sizeof(int)));
    *h_err = 0;
  }
  if (n <= cap_flags) return 0;
  if (d_flags) SDPA_HIP_CHECK(hipFree(d_flags));
  cap_flags = n;
  SDPA_HIP_CHECK(hipMalloc(&d_flags, sizeof(int) * cap_flags));
  SDPA_HIP_CHECK(hipMemset(d_flags, 0, sizeof(int) * cap_flags));
  epoch = 0;
  return 0;
}
int DenseWork::ensure_ll(size_t nwords) {
  if (nwords <= cap_ll) return 0;
  if (d_ll) SDPA_HIP_CHECK(hipFree(d_ll));
  cap_ll = nwords;
  SDPA_HIP_CHECK(hipMalloc(&d_ll, sizeof(unsigned long long) * cap_ll));
  SDPA_HIP_CHECK(hipMemset(d_ll, 0, sizeof(unsigned long long) * cap_ll));
  SDPA_HIP_CHECK(hipDeviceSynchronize());
  ll_epoch = 0;
  return 0;
}
int DenseWork::clear_err(hipStream_t s) {
  if (!d_err) return 0;
  SDPA_HIP_CHECK(hipMemsetAsync(d_err, 0, sizeof(int), s));
  SDPA_HIP_CHECK(hipStreamSynchronize(s));
  *h_err = 0;
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
  if (side) (void)hipStreamSynchronize(side);
  for (DescCopy& e : desc_pool) (void)hipFree(e.d);
  desc_pool.clear();
  for (auto e : evt0) (void)hipEventDestroy(e);
  for (auto e : evt1) (void)hipEventDestroy(e);
  for (auto e : evq0) (void)hipEventDestroy(e);
  for (auto e : evq1) (void)hipEventDestroy(e);
  evt0.clear();
  evt1.clear();
  evq0.clear();
  evq1.clear();
  evt_flops.clear();
  evq_flops.clear();
  trail_pending = panel_pending = 0;
  for (auto e : evp) (void)hipEventDestroy(e);
  for (auto e : evn) (void)hipEventDestroy(e);
  evp.clear();
  evn.clear();
  if (evjoin) (void)hipEventDestroy(evjoin);
  evjoin = nullptr;
  if (side) (void)hipStreamDestroy(side);
  side = nullptr;
  if (d_vec) (void)hipFree(d_vec);
  d_vec = nullptr; cap_vec = 0;
  if (d_flags) (void)hipFree(d_flags);
  if (d_err) (void)hipFree(d_err);
  if (h_err) (void)hipHostFree(h_err);
  d_flags = nullptr; d_err = nullptr; h_err = nullptr; cap_flags = 0; epoch = 0;
  if (d_ll) (void)hipFree(d_ll);
  d_ll = nullptr; cap_ll = 0; ll_max_nb = ll_max_nb32 = -1; ll_epoch = 0;
  if (d_ptag) (void)hipFree(d_ptag);
  d_ptag = nullptr; ptag_epoch = 0;
  if (d_dinv) (void)hipFree(d_dinv);
  d_dinv = nullptr;
  cap_dinv = 0;
}

int gemm(hipStream_t s, DenseWork* w, const double* A, int lda, const double* B, int ldb,
         double* C, int ldc, int M, int N, int K, int transA, int transB, double alpha,
         double beta, int lower_only) {
  if (M <= 0 || N <= 0) return 0;
  GemmItem it{A, B, C, M, N, K, lda, ldb, ldc};
  const void* di = nullptr;
  if (int e = w->desc_copy(s, &it, sizeof(it), &di)) return e;
  return gemm_batched(s, static_cast<const GemmItem*>(di), 1, M, N, transA, transB, alpha, beta,
                      lower_only);
}

// ---------------------------------------------------------------------------
// Cholesky of the kb x kb diagonal block at (k0, k0) of every matrix in the
// batch, and the inverse of that factor into dinv (64x64, column-major,
// strictly upper part zero) at slot (item, k0/64).  One 256-thread
// work-group per item, the block in LDS.
//
// Factor: four 16-wide mini-panels.  Wave 0 factors a mini-panel with its
// rows in registers (lane = row): the pivot comes by readlane, the column is
// scaled by rsqrt of the pivot (one transcendental per column on the serial
// chain instead of a square root and a divide), and the entries the rank-1
// update needs are read back from LDS as broadcasts.  All four waves then
// apply the mini-panel to the trailing block with 16x16x4 MFMAs.
// Inverse: the 16x16 diagonal inverses by right-looking substitution and
// the doubling W21 = -W22 L21 W11, first on 16x16 blocks, then on 32x32
// blocks (MFMAs); the waves idle during a mini-panel factor compute the
// inverses of the blocks already final.
// Pivot rules: mode 0 = dpotrf (d > 0 else fail), mode 1 = rATL_dpotrfL
// (d < -1e-6 fails, d < 1e-14 becomes 1e100).  Rows/columns beyond kb are
// padded with the identity.
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// One 16x16 tile C = A(16 x K) * B(K x 16) from LDS.  A is column-major
// (A[i][k] = A[i + lda k]); B[k][j] = B[k + ldb j], or B[j + ldb k] when
// TRB (B given as the rows of its transpose).  On return lane l, element v
// holds C[l % 16][4 v + l / 16].  K is a multiple of 8.
template <int TRB>
__device__ __forceinline__ v4d mma16_lds(const double* A, int lda, const double* B, int ldb,
                                         int K) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  v4d acc0 = v4d{0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
  for (int kk = 0; kk < K; kk += 8) {
    const int ka = kk + lk, kb2 = kk + 4 + lk;
    const double b0 = TRB ? B[li + ldb * ka] : B[ka + ldb * li];
    const double b1 = TRB ? B[li + ldb * kb2] : B[kb2 + ldb * li];
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, A[li + lda * ka], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, A[li + lda * kb2], acc1, 0, 0, 0);
  }
  return acc0 + acc1;
}

// Column j (lane % 16) of inv(L_oo), the 16x16 diagonal block of T at
// (o, o), by right-looking substitution (1 / L_ii = rdiag[o + i]), into W.
__device__ __forceinline__ void diag16_inverse(const double* T, const double* rdiag, double* W,
                                               int o) {
  const int lane = threadIdx.x & 63, j = lane & 15;
  double acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const double wi = acc[i] * rdiag[o + i];
    acc[i] = wi;
#pragma unroll
    for (int m = i + 1; m < 16; ++m) acc[m] -= T[(o + m) + TLD * (o + i)] * wi;
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) W[(o + i) + TLD * (o + j)] = (i >= j) ? acc[i] : 0.0;
  }
}

// W[h+1][h] = -W[h+1][h+1] (L[h+1][h] W[h][h]) on 16x16 blocks, one wave;
// Sw = 16x16 scratch (ld 17) private to the wave.
__device__ __forceinline__ void offdiag16_inverse(const double* T, double* W, double* Sw, int h) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const v4d u = mma16_lds<0>(&T[16 * (h + 1) + TLD * 16 * h], TLD, &W[16 * h + TLD * 16 * h],
                             TLD, 16);
#pragma unroll
  for (int v = 0; v < 4; ++v) Sw[li + 17 * (4 * v + lk)] = u[v];
  __builtin_amdgcn_wave_barrier();
  const v4d x = mma16_lds<0>(&W[16 * (h + 1) + TLD * 16 * (h + 1)], TLD, Sw, 17, 16);
#pragma unroll
  for (int v = 0; v < 4; ++v) W[(16 * (h + 1) + li) + TLD * (16 * h + 4 * v + lk)] = -x[v];
}

// LDS of the 64x64 diagonal-block factorization: T holds the block (its
// lower part becomes L), W receives inv(L) (it must be zero on entry: only
// the lower blocks are written), S/S1/S2 are scratch, rdiag the reciprocal
// pivots, failc the 1-based failing column (0 if none).
struct Chol64Lds {
  double T[TB * TLD];
  double W[TB * TLD];
  double S[32 * 33];
  double S1[16 * 17];
  double S2[16 * 17];
  double rdiag[TB];
  int failc;
};

// Hook of the 64x64 factorization: called by all threads with q once the
// columns 16q..16q+15 of the factor (in L.T) and the inverse of its 16x16
// diagonal tile (in L.W) are final and visible to the work-group.
struct NoPub {
  __device__ __forceinline__ void operator()(int) const {}
};

// Factor the block in L.T (256 threads; L.failc must be 0 and W zero, both
// visible to the work-group): 16-column mini-panels on wave 0, trailing
// 16x16 tiles on MFMA, and the first three diagonal 16x16 inverses and
// W[1][0] built alongside on waves 1 and 2.  MODE 0: LAPACK (a pivot that
// is not > 0 fails); MODE 1: SDPA's rule (sdpa_dpotrf.cpp:31-33: < -1e-6
// fails, < 1e-14 becomes 1e100).  kb: the block's true size (the padding
// beyond it is the identity).  pub(q) is called for q = 0, 1, 2 (the last
// one by chol64_inverse_finish).  Returns L.failc.
template <int MODE, class Pub = NoPub>
__device__ __forceinline__ int chol64_factor(Chol64Lds& L, int kb, Pub pub = Pub()) {
  double* T = L.T;
  double* W = L.W;
  double* S1 = L.S1;
  double* rdiag = L.rdiag;
  int& failc_s = L.failc;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  (void)t;
#pragma unroll 1
  for (int p = 0; p < 4; ++p) {
    const int c0 = 16 * p;
    if (wv == 0) {
      // lane = row.  Serial chain per column: pivot -> rsqrt -> the next
      // pivot, kept in uniform registers: d_{c+1} = a - l^2 with a the entry
      // (c+1, c+1) before column c's update and l = L[c+1][c] (the same fma
      // the vector update does at lane c+1).  The entries of column c for the
      // next two rows come by readlane; the later rows' multipliers are read
      // back from LDS as broadcasts, off the chain.  A failed pivot lets
      // NaN/Inf through; the block is then reported and discarded.
      const int r = lane;
      double v[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) v[jj] = T[r + TLD * (c0 + jj)];
      int failc = failc_s;
      double dcur = readlane_d(v[0], c0);
      double anext = readlane_d(v[1], c0 + 1);
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const int c = c0 + jj;
        const bool bad = (MODE == 0) ? !(dcur > 0.0) : (dcur < -1.0e-6);
        failc = (failc == 0 && c < kb && bad) ? c + 1 : failc;
        const double dd = (MODE == 1 && dcur < 1.0e-14) ? 1.0e+100 : dcur;
        double e1 = 0.0, e2 = 0.0;
        if (jj + 1 < 16) e1 = readlane_d(v[jj], c + 1);
        if (jj + 2 < 16) e2 = readlane_d(v[jj], c + 2);
        const double rs = rsqrt(dd);
        const double lrc = (r == c) ? dd * rs : v[jj] * rs;
        T[r + TLD * c] = lrc;
        if (r == 0) rdiag[c] = rs;
        double dnext = 0.0, anext2 = 0.0;
        if (jj + 1 < 16) {
          const double l1 = e1 * rs;
          v[jj + 1] = fma(-lrc, l1, v[jj + 1]);
          dnext = fma(-l1, l1, anext);
        }
        if (jj + 2 < 16) {
          const double l2 = e2 * rs;
          v[jj + 2] = fma(-lrc, l2, v[jj + 2]);
          anext2 = readlane_d(v[jj + 2], c + 2);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int kk = jj + 3; kk < 16; ++kk) v[kk] = fma(-lrc, T[(c0 + kk) + TLD * c], v[kk]);
        dcur = dnext;
        anext = anext2;
      }
      if (r == 0) failc_s = failc;
    } else if (wv == 1 && p >= 1) {
      diag16_inverse(T, rdiag, W, 16 * (p - 1));  // block p-1 is final
    } else if (wv == 2 && p == 3) {
      offdiag16_inverse(T, W, S1, 0);  // W[1][0] from D_0, D_1, L[1][0]
    }
    lds_barrier();
    if (p >= 1) pub(p - 1);
    // trailing update of the lower 16x16 tiles (bi, bj), p < bj <= bi < 4:
    // T[bi][bj] -= L[bi][p] L[bj][p]^T   (K = 16)
    const int nt = 3 - p, ntiles = nt * (nt + 1) / 2;
    for (int q = wv; q < ntiles; q += 4) {
      int bj = p + 1, rem = q;
      while (rem >= 4 - bj) { rem -= 4 - bj; ++bj; }
      const int bi = bj + rem;
      const v4d acc = mma16_lds<1>(&T[16 * bi + TLD * c0], TLD, &T[16 * bj + TLD * c0], TLD, 16);
#pragma unroll
      for (int v = 0; v < 4; ++v) T[(16 * bi + li) + TLD * (16 * bj + 4 * v + lk)] -= acc[v];
    }
    if (p < 3) lds_barrier();
  }
  return failc_s;
}

// Complete W = inv(L) of a block factored by chol64_factor (256 threads;
// T is only read, so the factor may be written out concurrently).  Ends
// with a barrier: W is complete for every thread.  pub(3) is called once the
// last 16x16 diagonal inverse is in W.
template <class Pub = NoPub>
__device__ __forceinline__ void chol64_inverse_finish(Chol64Lds& L, Pub pub = Pub()) {
  double* T = L.T;
  double* W = L.W;
  double* S = L.S;
  double* S2 = L.S2;
  double* rdiag = L.rdiag;
  constexpr int SLD = 33;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  (void)t;
  // D_3 (wave 1) alongside U = L[B][A] W[A][A] (A = 0:32, B = 32:64; waves
  // 0, 2, 3 share its four 16x16 tiles)
  if (wv == 1) {
    diag16_inverse(T, rdiag, W, 48);
  } else {
    for (int q = (wv == 0 ? 0 : wv - 1); q < 4; q += 3) {
      const int ti = q & 1, tj = q >> 1;
      const v4d u = mma16_lds<0>(&T[(32 + 16 * ti) + TLD * 0], TLD, &W[0 + TLD * (16 * tj)], TLD,
                                 32);
#pragma unroll
      for (int v = 0; v < 4; ++v) S[(16 * ti + li) + SLD * (16 * tj + 4 * v + lk)] = u[v];
    }
  }
  lds_barrier();
  pub(3);
  if (wv == 0) offdiag16_inverse(T, W, S2, 2);  // W[3][2] from D_2, D_3, L[3][2]
  lds_barrier();
  // W[B][A] = -W[B][B] U, wave = 16x16 output tile (ti, tj)
  {
    const int ti = wv & 1, tj = wv >> 1;
    const v4d x = mma16_lds<0>(&W[(32 + 16 * ti) + TLD * 32], TLD, &S[0 + SLD * (16 * tj)], SLD, 32);
#pragma unroll
    for (int v = 0; v < 4; ++v) W[(32 + 16 * ti + li) + TLD * (16 * tj + 4 * v + lk)] = -x[v];
  }
  lds_barrier();
}

// The diagonal-block work of one item (256 threads).  kp >= 0: first apply
// the update by the 64 columns [kp, kp + 64) of L to the block,
// A[k0.., k0..] -= L[k0.., kp..] L[k0.., kp..]^T (the block's share of the
// previous step's rest-of-panel update).
template <int MODE>
__device__ __forceinline__ void diag_block(const MatDesc& M, int b, int k0, int kp,
                                           int* __restrict__ info, double* __restrict__ dinv,
                                           int nslot) {
  const int kb = min(TB, M.n - k0);
  __shared__ Chol64Lds L;
  double* T = L.T;
  double* W = L.W;
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  double* A = M.ptr + k0 + (size_t)M.ld * k0;
  {
    // a thread's 16 loads in flight at once, then the LDS stores
    double a[16];
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = cq + 4 * q;
      a[q] = (r < kb && c < kb) ? A[r + (size_t)M.ld * c] : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) T[r + TLD * (cq + 4 * q)] = a[q];
    const double* X = M.ptr + k0 + (size_t)M.ld * (kp < 0 ? 0 : kp);
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = (kp >= 0 && r < kb) ? X[r + (size_t)M.ld * (cq + 4 * q)] : 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) W[r + TLD * (cq + 4 * q)] = a[q];
  }
  if (t == 0) L.failc = 0;
  lds_barrier();
  if (kp >= 0) {
    // T[bi][bj] -= X[bi] X[bj]^T on the ten lower 16x16 tiles (K = 64)
    for (int q = wv; q < 10; q += 4) {
      int bj = 0, rem = q;
      while (rem >= 4 - bj) { rem -= 4 - bj; ++bj; }
      const int bi = bj + rem;
      const v4d acc = mma16_lds<1>(&W[16 * bi], TLD, &W[16 * bj], TLD, 64);
#pragma unroll
      for (int v = 0; v < 4; ++v) T[(16 * bi + li) + TLD * (16 * bj + 4 * v + lk)] -= acc[v];
    }
    lds_barrier();
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) W[r + TLD * (cq + 4 * q)] = 0.0;
    lds_barrier();
  }
  const int failc = chol64_factor<MODE>(L, kb);
  if (failc) {
    if (t == 0) info[b] = k0 + failc;
    return;
  }
  // write the factor back (lower part of the block)
  {
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = cq + 4 * q;
      if (r >= c && r < kb && c < kb) A[r + (size_t)M.ld * c] = T[r + TLD * c];
    }
  }
  chol64_inverse_finish(L);
  double* Dv = dinv + ((size_t)b * nslot + k0 / TB) * TB * TB;
  for (int e = t; e < TB * TB; e += 256) Dv[e] = W[(e & 63) + TLD * (e >> 6)];
}


template <int MODE>
__global__ __launch_bounds__(256) void k_potrf_diag(const MatDesc* __restrict__ mats, int k0,
                                                    int* __restrict__ info,
                                                    double* __restrict__ dinv, int nslot) {
  const int b = blockIdx.x;
  const MatDesc M = mats[b];
  if (M.n <= k0) return;
  if (info[b] != 0) return;
  diag_block<MODE>(M, b, k0, -1, info, dinv, nslot);
}

// Rows below the diagonal block: L21 := A21 * inv(L11)^T, in place.  One
// work-group per 16 rows; wave w computes output columns [16 w, 16 w + 16)
// with K = 64, operands loaded straight into registers (A21 rows from the
// matrix, inv(L11) from dinv).  All waves finish their MFMAs (so every load
// of the 16 rows has been consumed) before any wave overwrites them.
// grid = (ceil(maxRows/16), nitems).
__global__ __launch_bounds__(256) void k_trsm_panel(const MatDesc* __restrict__ mats, int k0,
                                                    const int* __restrict__ info,
                                                    const double* __restrict__ dinv, int nslot) {
  const int b = blockIdx.y;
  const MatDesc M = mats[b];
  if (M.n <= k0 + TB) return;
  if (info[b] != 0) return;
  const int rows = M.n - k0 - TB;
  const int r0 = blockIdx.x * 16;
  if (r0 >= rows) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  double* A21 = M.ptr + (k0 + TB) + (size_t)M.ld * k0;
  const double* Dv = dinv + ((size_t)b * nslot + k0 / TB) * TB * TB;
  const double* pa = A21 + min(r0 + li, rows - 1) + (size_t)M.ld * lk;
  const double* pw = Dv + (16 * w + li) + TB * lk;
  double av[16], wv[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    av[s] = pa[(size_t)M.ld * 4 * s];
    wv[s] = pw[TB * 4 * s];
  }
  v4d acc0 = v4d{0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
  for (int s = 0; s < 16; s += 2) {
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(wv[s], av[s], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(wv[s + 1], av[s + 1], acc1, 0, 0, 0);
  }
  const v4d acc = acc0 + acc1;
  __syncthreads();
  if (r0 + li < rows) {
#pragma unroll
    for (int v = 0; v < 4; ++v) A21[(r0 + li) + (size_t)M.ld * (16 * w + 4 * v + lk)] = acc[v];
  }
}

// Update C[r, c] -= sum_{k < kw} L[r, k0 + k] L[c, k0 + k] for columns c in
// [cb, ce) and rows c <= r < n: the update of the rest of a panel after an
// inner 64-column step (kw = 64), or the trailing update of an outer panel
// (kw = 256), whole or split into the next panel's columns and the rest.
// One wave per (16 TM) x (16 TN) output tile, operands loaded straight from
// the matrix into registers (16 consecutive rows of one column per 16-lane
// group: four 128-byte segments per load), D k-steps of loads in flight; no
// LDS, no barrier.  Rows beyond the matrix are clamped on load and masked on
// store.  grid = (lower tiles, nitems): blockIdx.x enumerates the tiles
// column by column, rows from the diagonal down.
// One (16 TM) x (16 TN) output tile at (wm, wn) of the update
// C[cb + r][cb + c] -= sum_{k < kw} L[cb + r][k0 + k] L[cb + c][k0 + k], by one
// wave: operands loaded straight from the matrix into registers (16
// consecutive rows of one column per 16-lane group: four 128-byte segments
// per load), D k-steps of loads in flight, no LDS.  Rows beyond the matrix are
// clamped on load and masked on store; only r >= c is written.
template <int TM, int TN, int D>
__device__ __forceinline__ void syrk_tile(const MatDesc& M, int k0, int kw, int cb, int rows,
                                          int cend, int wm, int wn) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const size_t ld = M.ld;
  const double* P = M.ptr + cb + ld * k0;  // L[cb.., k0..]
  const double* pr[TM];
  const double* pc[TN];
#pragma unroll
  for (int a = 0; a < TM; ++a) pr[a] = P + min(wm + 16 * a + li, rows - 1) + ld * lk;
#pragma unroll
  for (int a = 0; a < TN; ++a) pc[a] = P + min(wn + 16 * a + li, cend - 1) + ld * lk;
  v4d acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int c = 0; c < TN; ++c) acc[a][c] = v4d{0.0, 0.0, 0.0, 0.0};
  double ra[D][TM], ca[D][TN];
  const int ns = kw >> 2;
  const size_t st = ld * 4;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const size_t off = st * (d < ns ? d : 0);
#pragma unroll
    for (int a = 0; a < TM; ++a) ra[d][a] = pr[a][off];
#pragma unroll
    for (int a = 0; a < TN; ++a) ca[d][a] = pc[a][off];
  }
  for (int s0 = 0; s0 < ns; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int s = s0 + d;
      if (s < ns) {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int c = 0; c < TN; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[d][c], ra[d][a], acc[a][c], 0, 0, 0);
        if (s + D < ns) {
          const size_t off = st * (s + D);
#pragma unroll
          for (int a = 0; a < TM; ++a) ra[d][a] = pr[a][off];
#pragma unroll
          for (int a = 0; a < TN; ++a) ca[d][a] = pc[a][off];
        }
      }
    }
  }
  double* C = M.ptr + cb + ld * cb;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int c = 0; c < TN; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gr = wm + 16 * a + li, gc = wn + 16 * c + 4 * v + lk;
        if (gr < rows && gc < cend && gr >= gc) C[gr + ld * gc] -= acc[a][c][v];
      }
}

// Tile g of the lower-triangular tile grid (R row tiles, Cn column tiles, a
// column tile as wide as r row tiles, so column j's tiles are rows i >= r j)
// in a blocked order: bands of sbc columns, each band a triangle at the
// diagonal (rows [r c0, r c0 + r bw)) and then full rows below; inside the
// triangle column by column, below it in row bands of r sbc rows, column by
// column inside a row band.  Consecutive tiles then share a few row and
// column slices of the panel.  Returns false past the last tile.
__host__ __device__ inline bool blocked_tile(int g, int R, int Cn, int r, int sbc, int& i,
                                             int& j) {
  for (int c0 = 0; c0 < Cn; c0 += sbc) {
    const int bw = min(sbc, Cn - c0), rs = r * c0;
    if (rs >= R) return false;
    const int rd = min(R, rs + r * bw);
    // the triangle: column jj holds rows [rs + r jj, rd)
    int tri = 0;
    if (rd - rs == r * bw) {
      tri = r * bw * (bw + 1) / 2;
    } else {
      for (int jj = 0; jj < bw; ++jj) tri += max(0, rd - rs - r * jj);
    }
    if (g < tri) {
      for (int jj = 0; jj < bw; ++jj) {
        const int cnt = max(0, rd - rs - r * jj);
        if (g < cnt) {
          i = rs + r * jj + g;
          j = c0 + jj;
          return true;
        }
        g -= cnt;
      }
    }
    g -= tri;
    const int below = (R - rd) * bw;
    if (g < below) {
      const int sbr = r * sbc;
      const int band = g / (sbr * bw), off = g - band * sbr * bw;
      const int r0 = rd + band * sbr, nr = min(sbr, R - r0);
      j = c0 + off / nr;
      i = r0 + off % nr;
      return true;
    }
    g -= below;
  }
  return false;
}

// Column band of the blocked tile order (0: column by column, the tiles of
// a bounded grid); SDPA_SYRK_BAND overrides it.
static int syrk_band() {
  static const int band = [] {
    const char* e = getenv("SDPA_SYRK_BAND");
    const int v = e ? atoi(e) : 8;
    return v >= 0 && v <= 64 ? v : 8;
  }();
  return band;
}

template <int TM, int TN, int D>
__global__ __launch_bounds__(64) void k_syrk_update(const MatDesc* __restrict__ mats, int k0,
                                                    int kw, int cb, int ce,
                                                    const int* __restrict__ info, int sbc) {
  constexpr int BMt = 16 * TM, BNt = 16 * TN;
  const int b = blockIdx.y;
  const MatDesc M = mats[b];
  if (info[b] != 0) return;
  const int rows = M.n - cb;
  const int cend = min(ce, M.n) - cb;
  if (rows <= 0 || cend <= 0) return;
  const int R = (rows + BMt - 1) / BMt, Cn = (cend + BNt - 1) / BNt;
  if (sbc > 0) {
    // one matrix, one tile per work-group: work-groups go to the 8 XCDs in
    // turn (work-group w to XCD w % 8), so XCD x is given the x-th eighth of
    // the blocked order, in its own dispatch order, and its L2 holds the
    // panel slices its concurrent tiles share
    static_assert(BNt % BMt == 0, "column tile a multiple of the row tile");
    const int per = gridDim.x >> 3;
    const int g = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    int i = 0, j = 0;
    if (!blocked_tile(g, R, Cn, BNt / BMt, sbc, i, j)) return;
    syrk_tile<TM, TN, D>(M, k0, kw, cb, rows, cend, i * BMt, j * BNt);
    return;
  }
  // a bounded grid walks the tiles (persistent waves); the tiles are
  // enumerated column by column, rows from the diagonal down
  for (int tile = blockIdx.x;; tile += gridDim.x) {
    int g = tile, i = 0, j = 0;
    for (; j < Cn; ++j) {
      const int cnt = R - (j * BNt) / BMt;
      if (g < cnt) { i = (j * BNt) / BMt + g; break; }
      g -= cnt;
    }
    if (j >= Cn) return;
    syrk_tile<TM, TN, D>(M, k0, kw, cb, rows, cend, i * BMt, j * BNt);
  }
}

// One step of a panel factorization, fused into one launch: work-group 0
// factors the diagonal block at k0 after applying to it the update by the
// previous 64 columns [kp, k0), kp = k0 - 64; the other work-groups (4 waves,
// one 32x32 tile each) apply that update to the rest of the panel, rows
// [k0, n) x columns [k0, pend) less the diagonal block.  The factorization's
// serial chain then waits on one launch per step for the update and the
// diagonal block together.  grid = (1 + ceil(tiles / 4), nitems).
template <int MODE>
__global__ __launch_bounds__(256, 2) void k_panel_step(const MatDesc* __restrict__ mats, int k0,
                                                    int pend, int* __restrict__ info,
                                                    double* __restrict__ dinv, int nslot) {
  const int b = blockIdx.y;
  const MatDesc M = mats[b];
  if (M.n <= k0) return;
  if (info[b] != 0) return;
  const int kp = k0 - TB;
  if (blockIdx.x == 0) {
    diag_block<MODE>(M, b, k0, kp, info, dinv, nslot);
    return;
  }
  // 32 x 64 tiles: row tile i, column tile j, from the diagonal down
  // (i >= 2 j); rows [0, 64) x columns [0, 64) is the diagonal block
  const int rows = M.n - k0, cend = min(pend, M.n) - k0;
  const int R = (rows + 31) / 32, Cn = (cend + 63) / 64;
  int g = (blockIdx.x - 1) * 4 + (threadIdx.x >> 6);
  for (int j = 0; j < Cn; ++j) {
    const int i0 = j == 0 ? 2 : 2 * j;
    const int cnt = R - i0;
    if (cnt <= 0) continue;
    if (g < cnt) {
      syrk_tile<2, 4, 3>(M, kp, TB, k0, rows, cend, (i0 + g) * 32, j * 64);
      return;
    }
    g -= cnt;
  }
}

static int panel_step_tiles(int maxn, int k0, int pend) {
  const int rows = maxn - k0, cend = std::min(pend, maxn) - k0;
  const int R = (rows + 31) / 32, Cn = (cend + 63) / 64;
  int n = 0;
  for (int j = 0; j < Cn; ++j) n += std::max(0, R - (j == 0 ? 2 : 2 * j));
  return n;
}

template <int TM, int TN, int D>
static void launch_syrk_t(hipStream_t s, const MatDesc* d_mats, const int* d_info, int nitems,
                          int rows, int cols, int k0, int kw, int cb, int ce, int maxwaves) {
  constexpr int BMt = 16 * TM, BNt = 16 * TN;
  const int R = (rows + BMt - 1) / BMt, Cn = (cols + BNt - 1) / BNt;
  int ntiles = 0;
  for (int j = 0; j < Cn; ++j) ntiles += R - (j * BNt) / BMt;
  if (maxwaves > 0 && nitems == 1) ntiles = std::min(ntiles, maxwaves);
  // one matrix and one tile per work-group: the XCD-aware blocked tile order
  const int sbc = (nitems == 1 && maxwaves <= 0) ? syrk_band() : 0;
  if (sbc) ntiles = (ntiles + 7) / 8 * 8;
  hipLaunchKernelGGL((k_syrk_update<TM, TN, D>), dim3(ntiles, nitems), dim3(64), 0, s, d_mats, k0,
                     kw, cb, ce, d_info, sbc);
}

// Tile shapes measured on MI355X (FP64): 32x64 tiles with 3 k-steps in flight
// for the wide trailing updates; 32x32 tiles (more waves, deeper prefetch)
// when the column range is at most one outer panel.  maxwaves > 0 bounds the
// grid (persistent waves), so an update running beside the panel
// factorization leaves no backlog of work-groups in the dispatcher.
static void launch_syrk(hipStream_t s, const MatDesc* d_mats, const int* d_info, int nitems,
                        int maxn, int k0, int kw, int cb, int ce, int maxwaves = 0) {
  const int rows = maxn - cb, cols = std::min(ce, maxn) - cb;
  if (rows <= 0 || cols <= 0) return;
  if (cols > 256)
    launch_syrk_t<2, 4, 3>(s, d_mats, d_info, nitems, rows, cols, k0, kw, cb, ce, maxwaves);
  else if (kw > 64)
    launch_syrk_t<2, 2, 8>(s, d_mats, d_info, nitems, rows, cols, k0, kw, cb, ce, maxwaves);
  else
    launch_syrk_t<2, 2, 4>(s, d_mats, d_info, nitems, rows, cols, k0, kw, cb, ce, maxwaves);
}

// ---------------------------------------------------------------------------
// One outer panel [K0, K0 + 256) of a single large matrix in ONE launch.
// Work-group b owns the 64 rows r0 = K0 + 64 b .. r0 + 63 of the panel, held
// in registers for the whole launch: wave w keeps rows r0 + 16 w + (lane %
// 16) and columns K0 + 4 s + lane / 16 (s = 0..63) in P[s] -- exactly the
// MFMA operand and result layout, so the solves and updates of the row
// block never leave the registers.  Step k (64 columns) of the panel:
//   - work-group k factors its (updated) 64x64 diagonal block in LDS (the
//     shared chol64 routines, SDPA's pivot rule in MODE 1); as each
//     16-column strip of the factor becomes final it publishes the strip
//     and the inverse D_q of the strip's 16x16 diagonal tile;
//   - every work-group b > k solves its block of column block k strip by
//     strip as the strips arrive (blocked forward substitution on MFMA:
//     X_q = X_q D_q^T, then X_c -= X_q L_cq^T for the later strips c), so
//     only the last strip's solve follows the diagonal factorization; it
//     then stores L_bk, publishes it if b < 4 (the rows of the panel's later
//     diagonal blocks), and updates its column blocks k < j <= min(b, 3):
//     A_bj -= L_bk L_jk^T.  Work-group k + 1 (the next diagonal block)
//     instead applies each solved strip to its own diagonal block at once and
//     defers the store and publication of L_bk to its factorization.
// Work-group b only ever waits for work-groups with a smaller index (k < b,
// j < b), which the dispatcher has started before it, so the launch needs no
// co-residency.  Values pass between work-groups as epoch-tagged 64-bit
// words (low and high halves of a double, each with the launch's epoch in
// the upper 32 bits, written and read with agent-scope atomics): a reader
// that sees its epoch in both words has the value, with no fence and no
// cache maintenance.  The chain from one diagonal block to the next is then
// a single work-group's: fetch W_k, solve its own 64 x 64, update its
// diagonal block, factor it -- without a kernel boundary.
// tw layout (8192-word blocks): block k (k < 4) holds the strips of step k's
// diagonal factor (strip q: rows 16q..63 of columns 16q..16q+15, element
// e = row + 64 col at words 2048 q + 2 e); block 21 + k the inverses D_q of
// its 16x16 diagonal tiles (element e = i + 16 j at words 512 q + 2 e, lower
// part only); block 4 + 4 j + k holds L_jk; block 20 is the control word
// (WG 0's go / skip).
constexpr int kPanelTagBlocks = 25;
constexpr int kTagSpin = 1 << 22;  // bounded waits: a timed-out wait sets *err

typedef unsigned long long u64;

__device__ __forceinline__ void tag_store(u64* dst, double x, unsigned epoch) {
  const u64 e = (u64)epoch << 32;
  __hip_atomic_store(dst, (u64)(unsigned)__double2loint(x) | e, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(dst + 1, (u64)(unsigned)__double2hiint(x) | e, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// The 64x64 block published at src (column-major element e = row + 64 col
// in words 2e, 2e + 1) into LDS T (ld TLD); 256 threads, 16 elements each in
// two rounds of 8.  A round issues its 16 loads at once; while some of its
// elements are not yet valid it reloads all of those together (one round trip
// per poll, not one per element).  LOWER: the block is lower triangular and
// only its lower part is published; the upper part is zero-filled here.
template <bool LOWER>
__device__ __forceinline__ void tag_fetch64(const u64* src, unsigned epoch, double* T, int* err) {
  const int t = threadIdx.x;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    u64 w[16];
    unsigned pend = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = t + 256 * (8 * h + q);
      if (!LOWER || (e & 63) >= (e >> 6)) {
        w[2 * q] = __hip_atomic_load(src + 2 * e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w[2 * q + 1] =
            __hip_atomic_load(src + 2 * e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        w[2 * q] = w[2 * q + 1] = (u64)epoch << 32;  // a zero, valid
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if ((unsigned)(w[2 * q] >> 32) != epoch || (unsigned)(w[2 * q + 1] >> 32) != epoch)
        pend |= 1u << q;
    int spins = 0;
    while (pend) {
      // after one timed-out wait anywhere the launch stops waiting (its
      // results are then discarded: the give-up flag is reported)
      if (++spins > kTagSpin ||
          ((spins & 1023) == 0 &&
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = t + 256 * (8 * h + q);
        if ((pend >> q) & 1u) {
          w[2 * q] = __hip_atomic_load(src + 2 * e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          w[2 * q + 1] =
              __hip_atomic_load(src + 2 * e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if ((unsigned)(w[2 * q] >> 32) == epoch && (unsigned)(w[2 * q + 1] >> 32) == epoch)
          pend &= ~(1u << q);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = t + 256 * (8 * h + q);
      T[(e & 63) + TLD * (e >> 6)] =
          __hiloint2double((int)(unsigned)w[2 * q + 1], (int)(unsigned)w[2 * q]);
    }
  }
}

// Strip q of a diagonal factor in T (rows 16q..63 of columns 16q..16q+15)
// and D_q, the inverse of its 16x16 diagonal tile in W at (16q, 16q) (lower
// part), published to the strip area sdst and the inverse area ddst.
__device__ __forceinline__ void strip_publish(int q, const double* T, const double* W, u64* sdst,
                                              u64* ddst, unsigned epoch) {
  const int t = threadIdx.x;
  u64* st = sdst + 2048 * q;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = t + 256 * k, r = e & 63, c = e >> 6;
    if (r >= 16 * q) tag_store(st + 2 * e, T[r + TLD * (16 * q + c)], epoch);
  }
  const int i = t & 15, j = t >> 4;
  if (i >= j) tag_store(ddst + 512 * q + 2 * t, W[(16 * q + i) + TLD * (16 * q + j)], epoch);
}

// Strip Q and D_Q (as published by strip_publish) into LDS: the strip at its
// place in T, D_Q in Dq (16 x 16, ld 17, zero above the diagonal).  Five
// elements per thread, all loads in flight at once, pending ones re-polled
// together.
template <int Q>
__device__ __forceinline__ void strip_fetch(const u64* ssrc, const u64* dsrc, unsigned epoch,
                                            double* T, double* Dq, int* err) {
  const int t = threadIdx.x;
  u64 w[10];
  unsigned pend = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const bool need = k < 4 ? ((t & 63) >= 16 * Q) : ((t & 15) >= (t >> 4));
    const u64* a = k < 4 ? ssrc + 2048 * Q + 2 * (t + 256 * k) : dsrc + 512 * Q + 2 * t;
    if (need) {
      w[2 * k] = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      w[2 * k + 1] = __hip_atomic_load(a + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((unsigned)(w[2 * k] >> 32) != epoch || (unsigned)(w[2 * k + 1] >> 32) != epoch)
        pend |= 1u << k;
    } else {
      w[2 * k] = w[2 * k + 1] = (u64)epoch << 32;  // a zero, valid
    }
  }
  int spins = 0;
  while (pend) {
    if (++spins > kTagSpin ||
        ((spins & 1023) == 0 &&
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const u64* a = k < 4 ? ssrc + 2048 * Q + 2 * (t + 256 * k) : dsrc + 512 * Q + 2 * t;
      if ((pend >> k) & 1u) {
        w[2 * k] = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w[2 * k + 1] = __hip_atomic_load(a + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if ((unsigned)(w[2 * k] >> 32) == epoch && (unsigned)(w[2 * k + 1] >> 32) == epoch)
        pend &= ~(1u << k);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = t + 256 * k, r = e & 63, c = e >> 6;
    if (r >= 16 * Q)
      T[r + TLD * (16 * Q + c)] =
          __hiloint2double((int)(unsigned)w[2 * k + 1], (int)(unsigned)w[2 * k]);
  }
  Dq[(t & 15) + 17 * (t >> 4)] = __hiloint2double((int)(unsigned)w[9], (int)(unsigned)w[8]);
}

// acc[c] = sum over the 64 columns of register block KB of this wave's rows
// times the rows 16 c + (lane % 16) of the 64x64 LDS operand X:
// acc[c][v] = sum_t P_row[t] X[16 c + 4 v + lane / 16][t], i.e. the 16 x 64
// block (rows of the wave) x X^T in the register layout of P.
template <int KB>
__device__ __forceinline__ void rows_times_xt(const double (&P)[64], const double* X, v4d acc[4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 16; ++s) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[(16 * c + li) + TLD * (4 * s + lk)],
                                                    P[16 * KB + s], acc[c], 0, 0, 0);
  }
}

// acc[c] += the same product (no zeroing).
template <int KB>
__device__ __forceinline__ void rows_times_xt_acc(const double (&P)[64], const double* X,
                                                  v4d acc[4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[(16 * c + li) + TLD * (4 * s + lk)],
                                                    P[16 * KB + s], acc[c], 0, 0, 0);
  }
}

// The register block KB of the wave's rows into LDS T (the 64 x 64 block of
// the work-group's rows, column-major, ld TLD).
template <int KB>
__device__ __forceinline__ void rows_to_lds(const double (&P)[64], double* T) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int s = 0; s < 16; ++s) T[(16 * w + li) + TLD * (4 * s + lk)] = P[16 * KB + s];
}

template <int KB>
__device__ __forceinline__ void rows_store(const double (&P)[64], double* A, int ld, int r0, int m,
                                           int K0) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int row = r0 + 16 * w + li;
  if (row >= m) return;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int col = K0 + 64 * KB + 4 * s + lk;
    if (col < m) A[row + (size_t)ld * col] = P[16 * KB + s];
  }
}

// The next diagonal work-group's own block, A_bb -= X_q X_q^T, for the solved
// strip q of block K (all 64 rows of it in U, this wave's 16 in registers).
template <int K, int q>
__device__ __forceinline__ void own_strip_update(double (&P)[64], const double* U) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(U[(16 * c + li) + TLD * (16 * q + 4 * s4 + lk)],
                                                 P[16 * K + 4 * q + s4], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) P[16 * (K + 1) + 4 * c + v] -= acc[v];
  }
}

// Strip Q of step K for a work-group below the diagonal block: fetch the
// strip and D_Q, apply the previous strip to the work-group's own diagonal
// block (CHAIN: the next diagonal work-group), solve X_Q = X_Q D_Q^T (X_Q:
// the wave's 16 rows of columns 16Q..16Q+15 of block K, registers
// P[16 K + 4 Q + s]), keep it in U (CHAIN) and take it off the later strips:
// X_c -= X_Q L_cQ^T.
template <int K, int Q>
__device__ __forceinline__ void strip_step(double (&P)[64], Chol64Lds& L, double* U, double* Dq,
                                           bool chain, const u64* ssrc, const u64* dsrc,
                                           u64* xdst, unsigned epoch, int* err) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  strip_fetch<Q>(ssrc, dsrc, epoch, L.T, Dq + 272 * Q, err);
  lds_barrier();
  if constexpr (Q > 0) {
    if (chain) own_strip_update<K, Q - 1>(P, U);
  }
  {
    const double* D = Dq + 272 * Q;
    v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(D[li + 17 * (4 * s4 + lk)], P[16 * K + 4 * Q + s4],
                                                 acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) P[16 * K + 4 * Q + v] = acc[v];
  }
  if (chain) {
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      U[(16 * w + li) + TLD * (16 * Q + 4 * s4 + lk)] = P[16 * K + 4 * Q + s4];
  }
  if (xdst) {
    // the solved strip for the work-groups below that update with it (the
    // column layout of the whole 64 x 64 block L_bK: element row + 64 col)
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      tag_store(xdst + 2 * ((16 * w + li) + 64 * (16 * Q + 4 * s4 + lk)), P[16 * K + 4 * Q + s4],
                epoch);
  }
#pragma unroll
  for (int c = Q + 1; c < 4; ++c) {
    v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L.T[(16 * c + li) + TLD * (16 * Q + 4 * s4 + lk)],
                                                 P[16 * K + 4 * Q + s4], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) P[16 * K + 4 * c + v] -= acc[v];
  }
}

// Step K of the panel for a work-group below the diagonal block (b > K):
// L_bK = A_bK L_KK^-T strip by strip (strip_step).  The next diagonal
// work-group (b == K + 1 < nsteps) is the factorization's serial chain: it
// applies each solved strip to its own diagonal block (its only later column
// block in the panel) and goes straight on to factor it; its solved block
// L_bK is left in U, and the matrix store of it is made by waves 1-3 while
// wave 0 factors (panel_step_diag).  Work-groups b < 4 publish each solved
// strip of L_bK at once, for the work-groups below that update with it.  The work-group after it
// (b == K + 2) leaves the update of its own diagonal block by step K to the
// start of step K + 1, so that it is ready for step K + 1's first strips as
// soon as its column block K + 1 is up to date.
template <int K, int MODE>
__device__ __forceinline__ void panel_step_rows(double (&P)[64], Chol64Lds& L, double* U,
                                                double* Dq, int b, int nsteps, double* A, int ld,
                                                int r0, int m, int K0, const u64* tw, u64* twm,
                                                unsigned epoch, int* err) {
  const bool chain = K + 1 < 4 && b == K + 1 && b < nsteps;
  const bool defer_own = K + 2 < 4 && b == K + 2 && b < nsteps;
  if constexpr (K >= 1) {
    if (chain) {
      // the update of the own diagonal block by step K - 1, deferred from there
      rows_to_lds<K - 1>(P, U);
      lds_barrier();
      v4d acc[4];
      rows_times_xt<K - 1>(P, U, acc);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) P[16 * (K + 1) + 4 * c + v] -= acc[c][v];
    }
  }
  const u64* ssrc = tw + (size_t)K * 8192;
  const u64* dsrc = tw + (size_t)(21 + K) * 8192;
  // the panel's later diagonal blocks' rows are published strip by strip
  u64* xdst = (b < 4 && b < nsteps) ? twm + (size_t)(4 + 4 * b + K) * 8192 : nullptr;
  strip_step<K, 0>(P, L, U, Dq, chain, ssrc, dsrc, xdst, epoch, err);
  strip_step<K, 1>(P, L, U, Dq, chain, ssrc, dsrc, xdst, epoch, err);
  strip_step<K, 2>(P, L, U, Dq, chain, ssrc, dsrc, xdst, epoch, err);
  strip_step<K, 3>(P, L, U, Dq, chain, ssrc, dsrc, xdst, epoch, err);
  if (chain) {
    lds_barrier();
    own_strip_update<K, 3>(P, U);
    return;
  }
  rows_store<K>(P, A, ld, r0, m, K0);
  // A_bj -= L_bK L_jK^T for the panel's later column blocks j <= b
#pragma unroll
  for (int j = K + 1; j < 4; ++j) {
    if (j < nsteps && j <= b && !(defer_own && j == b)) {
      lds_barrier();  // the LDS operand of the previous product is consumed
      if (j < b) tag_fetch64<false>(tw + (size_t)(4 + 4 * j + K) * 8192, epoch, L.T, err);
      else rows_to_lds<K>(P, L.T);
      lds_barrier();
      v4d acc[4];
      rows_times_xt<K>(P, L.T, acc);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) P[16 * j + 4 * c + v] -= acc[c][v];
    }
  }
  lds_barrier();
}

// Step K for work-group K itself: factor the diagonal block of its rows,
// publishing the factor's strips and their tile inverses as they become
// final.  For K > 0 the work-group's solved block of the previous step,
// L_{K,K-1}, waits in U: waves 1-3 store it to the matrix while wave 0 runs
// the factorization's first column chain (it was published strip by strip
// as it was solved).  L.W is zero on entry (cleared at the
// start of the launch; no other step writes it).
template <int K, int MODE>
__device__ __forceinline__ void panel_step_diag(double (&P)[64], Chol64Lds& L, const double* U,
                                                double* A, int ld, int r0, int m, int K0,
                                                int* info, double* dinv, u64* twm,
                                                unsigned epoch) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, lk = lane >> 4;
  const int kb = min(TB, m - r0);
  // the block (identity beyond kb) into T
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int r = 16 * w + li, c = 4 * s + lk;
    L.T[r + TLD * c] = (r < kb && c < kb) ? P[16 * K + s] : (r == c ? 1.0 : 0.0);
  }
  if (t == 0) L.failc = 0;
  lds_barrier();
  if (K > 0 && w > 0) {
    const int col0 = K0 + 64 * (K > 0 ? K - 1 : 0);
    for (int e = t - 64; e < TB * TB; e += 192) {
      const int r = e & 63, c = e >> 6;
      if (r0 + r < m && col0 + c < m) A[(r0 + r) + (size_t)ld * (col0 + c)] = U[r + TLD * c];
    }
  }
  u64* sdst = twm + (size_t)K * 8192;
  u64* ddst = twm + (size_t)(21 + K) * 8192;
  const double* T = L.T;
  const double* W = L.W;
  auto pub = [=](int q) { strip_publish(q, T, W, sdst, ddst, epoch); };
  const int failc = chol64_factor<MODE>(L, kb, pub);
  if (failc && t == 0) atomicCAS(info, 0, r0 + failc);
  chol64_inverse_finish(L, pub);
  // the factor (lower part) to the matrix, the inverse to dinv
  {
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = cq + 4 * q;
      if (r >= c && r < kb && c < kb) A[(r0 + r) + (size_t)ld * (r0 + c)] = L.T[r + TLD * c];
    }
  }
  double* Dv = dinv + (size_t)(r0 / TB) * TB * TB;
  for (int e = t; e < TB * TB; e += 256) Dv[e] = L.W[(e & 63) + TLD * (e >> 6)];
}

// The previous panel's update of one work-group's rows (the left-looking
// prologue): A_bj -= L_b,prev L_j,prev^T for the column blocks j <= b of the
// panel, K = the previous panel's 256 columns [Kp, Kp + 256) in 64-column
// chunks; L_b,prev sits in registers (Y, the layout of P), each 64 x 64
// chunk of L_j,prev is staged in LDS, the next chunk loaded while the current
// one is multiplied.
__device__ __forceinline__ void chunk_load(const double* A, int ld, int m, int row0, int col0,
                                           double (&pre)[16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = t + 256 * q, r = row0 + (e & 63), c = col0 + (e >> 6);
    pre[q] = r < m ? A[r + (size_t)ld * c] : 0.0;
  }
}

__device__ __forceinline__ void chunk_to_lds(const double (&pre)[16], double* T) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = t + 256 * q;
    T[(e & 63) + TLD * (e >> 6)] = pre[q];
  }
}

__device__ __forceinline__ void panel_prologue(double (&P)[64], Chol64Lds& L, const double* A,
                                               int ld, int m, int K0, int Kp, int r0, int b,
                                               int nsteps) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int nj = min(b + 1, nsteps);  // column blocks below the diagonal
  double Y[64];
  {
    const int row = r0 + 16 * w + li;
    const bool rin = row < m;
    const double* src = A + (rin ? row : 0) + (size_t)ld * Kp;
#pragma unroll
    for (int s2 = 0; s2 < 64; ++s2) Y[s2] = rin ? src[(size_t)ld * (4 * s2 + lk)] : 0.0;
  }
  double pre[16];
  chunk_load(A, ld, m, K0, Kp, pre);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j < nj) {
      v4d acc[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        lds_barrier();  // the previous chunk is consumed
        chunk_to_lds(pre, L.T);
        lds_barrier();
        // the next chunk: (j, q + 1), or (j + 1, 0)
        const int nq = q < 3 ? q + 1 : 0, njj = q < 3 ? j : j + 1;
        if (njj < nj) chunk_load(A, ld, m, K0 + 64 * njj, Kp + 64 * nq, pre);
        if (q == 0) rows_times_xt_acc<0>(Y, L.T, acc);
        if (q == 1) rows_times_xt_acc<1>(Y, L.T, acc);
        if (q == 2) rows_times_xt_acc<2>(Y, L.T, acc);
        if (q == 3) rows_times_xt_acc<3>(Y, L.T, acc);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) P[16 * j + 4 * c + v] -= acc[c][v];
    }
  }
  lds_barrier();
}

template <int MODE>
__global__ __launch_bounds__(256) void k_panel_rows(double* __restrict__ A, int ld, int m,
                                                       int K0, int Kp, int* __restrict__ info,
                                                       double* __restrict__ dinv,
                                                       u64* __restrict__ tw, unsigned epoch,
                                                       int* __restrict__ err) {
  __shared__ Chol64Lds L;
  __shared__ double U[TB * TLD];
  __shared__ double Dq[4 * 272];
  __shared__ int skip_s;
  const int b = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, lk = lane >> 4;
  const int r0 = K0 + 64 * b;
  const int ncols = min(256, m - K0);
  const int nsteps = (ncols + 63) / 64;
  // go / skip: work-group 0 reads the factorization's info once (an earlier
  // launch may have failed) and publishes the decision; the others take it
  // from there, so all of them see the same decision
  u64* ctrl = tw + (size_t)20 * 8192;
  if (t == 0) {
    int skip = 0;
    if (b == 0) {
      skip = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
      __hip_atomic_store(ctrl, (u64)(unsigned)skip | ((u64)epoch << 32), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      u64 c = __hip_atomic_load(ctrl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int spins = 0;
      while ((unsigned)(c >> 32) != epoch) {
        if (++spins > kTagSpin) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          c = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        c = __hip_atomic_load(ctrl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      skip = (int)(c & 1);
    }
    skip_s = skip;
  }
  // the row block into registers (rows beyond m and columns beyond the
  // matrix read as zero)
  double P[64];
  {
    const int row = r0 + 16 * w + li;
    const bool rin = row < m;
    const double* src = A + (rin ? row : 0) + (size_t)ld * K0;
#pragma unroll
    for (int s = 0; s < 64; ++s) {
      const int col = 4 * s + lk;
      P[s] = (rin && col < ncols) ? src[(size_t)ld * col] : 0.0;
    }
  }
  // the diagonal step's inverse buffer starts at zero (only its factorization writes it)
  for (int e = t; e < TB * TB; e += 256) L.W[(e & 63) + TLD * (e >> 6)] = 0.0;
  lds_barrier();
  if (skip_s) return;
  if (Kp >= 0) panel_prologue(P, L, A, ld, m, K0, Kp, r0, b, nsteps);
#define PANEL_STEP(K)                                                                      \
  if (K < nsteps) {                                                                        \
    if (b == K) {                                                                          \
      /* the work-group's last step: its later column blocks are above the diagonal */    \
      panel_step_diag<K, MODE>(P, L, U, A, ld, r0, m, K0, info, dinv, tw, epoch);           \
      return;                                                                              \
    }                                                                                      \
    panel_step_rows<K, MODE>(P, L, U, Dq, b, nsteps, A, ld, r0, m, K0, tw, tw, epoch, err);     \
  }
  PANEL_STEP(0)
  PANEL_STEP(1)
  PANEL_STEP(2)
  PANEL_STEP(3)
#undef PANEL_STEP
}

int DenseWork::ensure_panel_tags() {
  if (d_ptag) return 0;
  SDPA_HIP_CHECK(hipMalloc(&d_ptag, sizeof(unsigned long long) * kPanelTagBlocks * 8192));
  SDPA_HIP_CHECK(hipMemset(d_ptag, 0, sizeof(unsigned long long) * kPanelTagBlocks * 8192));
  SDPA_HIP_CHECK(hipDeviceSynchronize());
  ptag_epoch = 0;
  return 0;
}

// Factor the outer panel [K0, K0 + 256) of the single m x m matrix A by one
// launch of k_panel_rows, after applying the previous panel [Kp, Kp + 256)
// to it (Kp < 0: none).  A's 64x64 diagonal inverses go to dinv, slot
// k0 / 64.
static int factor_panel_rows(hipStream_t s, DenseWork* w, double* A, int ld, int m, int K0,
                             int Kp, int mode, int* d_info, double* dinv) {
  if (int e = w->ensure_panel_tags()) return e;
  if (int e = w->ensure_flags(0)) return e;  // the give-up flag
  if (++w->ptag_epoch == 0) ++w->ptag_epoch;
  const int R = (m - K0 + 63) / 64;
  if (mode == 0)
    hipLaunchKernelGGL(k_panel_rows<0>, dim3(R), dim3(256), 0, s, A, ld, m, K0, Kp, d_info,
                       dinv, w->d_ptag, w->ptag_epoch, w->d_err);
  else
    hipLaunchKernelGGL(k_panel_rows<1>, dim3(R), dim3(256), 0, s, A, ld, m, K0, Kp, d_info,
                       dinv, w->d_ptag, w->ptag_epoch, w->d_err);
  return 0;
}

// Factor the outer panel [K0, K0 + PW): 64-column steps.  Step 0 is the
// diagonal block alone; every later step is one fused launch (the previous
// step's rest-of-panel update with the next diagonal block); each step ends
// with the rows below its diagonal block.
static void factor_panel(hipStream_t s, const MatDesc* d_mats, int nitems, int maxn, int K0,
                         int PW, int mode, int* d_info, double* dinv, int nslot) {
  const int pend = std::min(K0 + PW, maxn);
  for (int k0 = K0; k0 < pend; k0 += TB) {
    if (k0 == K0) {
      if (mode == 0)
        hipLaunchKernelGGL(k_potrf_diag<0>, dim3(nitems), dim3(256), 0, s, d_mats, k0, d_info,
                           dinv, nslot);
      else
        hipLaunchKernelGGL(k_potrf_diag<1>, dim3(nitems), dim3(256), 0, s, d_mats, k0, d_info,
                           dinv, nslot);
    } else {
      const int nwg = 1 + (panel_step_tiles(maxn, k0, K0 + PW) + 3) / 4;
      if (mode == 0)
        hipLaunchKernelGGL(k_panel_step<0>, dim3(nwg, nitems), dim3(256), 0, s, d_mats, k0,
                           K0 + PW, d_info, dinv, nslot);
      else
        hipLaunchKernelGGL(k_panel_step<1>, dim3(nwg, nitems), dim3(256), 0, s, d_mats, k0,
                           K0 + PW, d_info, dinv, nslot);
    }
    const int rows = maxn - k0 - TB;
    if (rows <= 0) break;
    hipLaunchKernelGGL(k_trsm_panel, dim3((rows + 15) / 16, nitems), dim3(256), 0, s, d_mats,
                       k0, d_info, dinv, nslot);
  }
}

// The side stream runs the wide trailing updates.  Its CU mask (the first
// three quarters of the mask bits) leaves the last quarter of the CUs to the
// panel factorization on the main stream: a trailing update with thousands
// of waves would otherwise hold every CU, and the one-work-group diagonal
// kernel that the whole factorization waits on could not start until it
// drained.  Measured on MI355X: with a contiguous mask and a full grid (a
// backlog of waves on the masked CUs) kernels of another stream run at full
// speed on the remaining CUs; masks interleaved at finer grain, or a grid
// that leaves masked CUs partly free, let them queue behind or share CUs with
// the update.
int DenseWork::ensure_side(int npanels) {
  if (!side) {
    int dev = 0, ncu = 0;
    SDPA_HIP_CHECK(hipGetDevice(&dev));
    SDPA_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu >= 128) {
      // SDPA_SIDE_CUS overrides the side stream's CU count (1 .. ncu - 1)
      int nside = ncu * 3 / 4;
      if (const char* env = getenv("SDPA_SIDE_CUS")) {
        const int v = atoi(env);
        if (v >= 1 && v < ncu) nside = v;
      }
      std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
      for (int i = 0; i < nside; ++i) mask[i >> 5] |= 1u << (i & 31);
      if (hipExtStreamCreateWithCUMask(&side, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
        (void)hipGetLastError();
        side = nullptr;
      }
    }
    if (!side) SDPA_HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  }
  // the events only order the two streams of this device: a device-scope
  // release (no system-scope cache writeback) is enough
  constexpr unsigned kEvFlags = hipEventDisableTiming | hipEventReleaseToDevice;
  if (!evjoin) SDPA_HIP_CHECK(hipEventCreateWithFlags(&evjoin, kEvFlags));
  while ((int)evp.size() < npanels) {
    hipEvent_t a, b;
    SDPA_HIP_CHECK(hipEventCreateWithFlags(&a, kEvFlags));
    SDPA_HIP_CHECK(hipEventCreateWithFlags(&b, kEvFlags));
    evp.push_back(a);
    evn.push_back(b);
  }
  return 0;
}

int DenseWork::ensure_trailing_events(int n) {
  while ((int)evt0.size() < n) {
    hipEvent_t a, b;
    SDPA_HIP_CHECK(hipEventCreate(&a));
    SDPA_HIP_CHECK(hipEventCreate(&b));
    evt0.push_back(a);
    evt1.push_back(b);
    evt_flops.push_back(0.0);
  }
  return 0;
}

int DenseWork::ensure_panel_events(int n) {
  while ((int)evq0.size() < n) {
    hipEvent_t a, b;
    SDPA_HIP_CHECK(hipEventCreate(&a));
    SDPA_HIP_CHECK(hipEventCreate(&b));
    evq0.push_back(a);
    evq1.push_back(b);
    evq_flops.push_back(0.0);
  }
  return 0;
}

int DenseWork::collect_trailing() {
  for (int q = 0; q < trail_pending; ++q) {
    float ms = 0.f;
    SDPA_HIP_CHECK(hipEventElapsedTime(&ms, evt0[q], evt1[q]));
    trail_ms += ms;
    trail_flops += evt_flops[q];
    ++trail_launches;
  }
  trail_pending = 0;
  for (int q = 0; q < panel_pending; ++q) {
    float ms = 0.f;
    SDPA_HIP_CHECK(hipEventElapsedTime(&ms, evq0[q], evq1[q]));
    panel_ms += ms;
    panel_flops += evq_flops[q];
    ++panel_launches;
  }
  panel_pending = 0;
  return 0;
}

int potrf_batched(hipStream_t s, DenseWork* w, const MatDesc* h_items, int nitems, int mode,
                  int* d_info, double* dinv_out) {
  if (nitems <= 0) return 0;
  int maxn = 0;
  for (int b = 0; b < nitems; ++b) maxn = h_items[b].n > maxn ? h_items[b].n : maxn;
  const int nslot = (maxn + TB - 1) / TB;
  double* dinv = dinv_out;
  if (!dinv) {
    if (int e = w->ensure_dinv((size_t)nitems * nslot * TB * TB)) return e;
    dinv = w->d_dinv;
  }
  const void* dm = nullptr;
  if (int e = w->desc_copy(s, h_items, sizeof(MatDesc) * nitems, &dm)) return e;
  const MatDesc* d_mats = static_cast<const MatDesc*>(dm);
  SDPA_HIP_CHECK(hipMemsetAsync(d_info, 0, sizeof(int) * nitems, s));
  // outer panels of PW = 256 columns; inside, 64-column steps whose updates
  // stay within the panel; then the trailing update with K = 256.
  constexpr int PW = 256;
  const int np = (maxn + PW - 1) / PW;
  if (nitems > 1 || np <= 2) {
    for (int K0 = 0; K0 < maxn; K0 += PW) {
      factor_panel(s, d_mats, nitems, maxn, K0, PW, mode, d_info, dinv, nslot);
      launch_syrk(s, d_mats, d_info, nitems, maxn, K0, PW, K0 + PW, maxn);
    }
  } else {
    // Look-ahead.  Panel p's trailing update is split by columns: panel
    // p+1's is the prologue of panel p+1's own launch on `s` (no separate
    // launch, no cross-stream wait in front of the next panel), panel p+2's
    // and the rest run on the CU-masked side stream beside the next panels.
    // Every update of a column block is ordered after the earlier ones that
    // write it; the side launches of panel p read the panel only after it is
    // factored (evp[p]).
    if (int e = w->ensure_side(np)) return e;
    SDPA_HIP_CHECK(hipEventRecord(w->evjoin, s));
    SDPA_HIP_CHECK(hipStreamWaitEvent(w->side, w->evjoin, 0));
    for (int p = 0; p < np; ++p) {
      const int K0 = p * PW;
      // panel p's columns carry the updates of panels 0 .. p-2 once the side
      // stream has run panel p-2's launch for them (evn[p-2], which the side
      // stream's order puts after every earlier side launch); panel p-1's
      // update is the panel kernel's own prologue
      if (p >= 2) SDPA_HIP_CHECK(hipStreamWaitEvent(s, w->evn[p - 2], 0));
      if (w->time_trailing) {
        if (int e = w->ensure_panel_events(w->panel_pending + 1)) return e;
        SDPA_HIP_CHECK(hipEventRecord(w->evq0[w->panel_pending], s));
      }
      if (int e = factor_panel_rows(s, w, h_items[0].ptr, h_items[0].ld, maxn, K0,
                                    p > 0 ? K0 - PW : -1, mode, d_info, dinv))
        return e;
      if (w->time_trailing) {
        SDPA_HIP_CHECK(hipEventRecord(w->evq1[w->panel_pending], s));
        // algorithmic flops of the launch: the previous panel's update of this
        // panel's lower part (2 K per element, K = 256) and the panel's own
        // Cholesky (w^2 h - w^3 / 3 for an h x w panel)
        const double h = maxn - K0, wc = std::min(PW, maxn - K0);
        const double elems = wc * h - wc * (wc - 1.0) / 2.0;
        w->evq_flops[w->panel_pending] = (p > 0 ? 2.0 * PW * elems : 0.0) + wc * wc * h -
                                         wc * wc * wc / 3.0;
        ++w->panel_pending;
      }
      if (K0 + PW >= maxn) break;
      SDPA_HIP_CHECK(hipEventRecord(w->evp[p], s));
      // the side stream, once panel p is factored: the panel-after-next's
      // columns, then the rest
      SDPA_HIP_CHECK(hipStreamWaitEvent(w->side, w->evp[p], 0));
      launch_syrk(w->side, d_mats, d_info, 1, maxn, K0, PW, K0 + 2 * PW, K0 + 3 * PW,
                  w->side_waves);
      SDPA_HIP_CHECK(hipEventRecord(w->evn[p], w->side));
      const int nrest = maxn - (K0 + 3 * PW);
      // timed: the launches that take the wide-tile kernel (more than 256 columns)
      const bool timed = w->time_trailing && nrest > 256;
      if (timed) {
        if (int e = w->ensure_trailing_events(w->trail_pending + 1)) return e;
        SDPA_HIP_CHECK(hipEventRecord(w->evt0[w->trail_pending], w->side));
      }
      launch_syrk(w->side, d_mats, d_info, 1, maxn, K0, PW, K0 + 3 * PW, maxn, w->side_waves);
      if (timed) {
        SDPA_HIP_CHECK(hipEventRecord(w->evt1[w->trail_pending], w->side));
        // algorithmic flops: 2 K per element of the lower triangle updated
        w->evt_flops[w->trail_pending] = (double)PW * nrest * (nrest + 1.0);
        ++w->trail_pending;
      }
    }
    SDPA_HIP_CHECK(hipEventRecord(w->evjoin, w->side));
    SDPA_HIP_CHECK(hipStreamWaitEvent(s, w->evjoin, 0));
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Mixed-precision factorization of the Schur matrix.  B (FP64, lower
// triangle) is left intact for the refinement residuals.  A32 = S B S in
// FP32, S = diag(B_ii^-1/2) (column equilibration: unit diagonal), is
// factored in place by 256-column panels: each panel is widened to FP64 and
// factored by the single-launch FP64 panel kernel (diagonal blocks, pivot
// rule, 64x64 inverses), written back as FP32, and applied to the trailing matrix by an
// FP32-MFMA update.  The 256x256 diagonal blocks of the FP64 panel factors
// are kept (D64) for the block inverses of the solves.
typedef float v4f __attribute__((ext_vector_type(4)));

// s_i = 1/sqrt(B_ii); a diagonal that is not positive sets *bad.  *bmax:
// max B_ii (bits of a non-negative double, which order like the integers),
// which bounds every |B_ij| of a positive definite B.
__global__ __launch_bounds__(256) void k_eq_scale(const double* __restrict__ B, int ld, int m,
                                                  double* __restrict__ scale, int* bad,
                                                  unsigned long long* bmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double d = 0.0;
  if (i < m) {
    d = B[i + (size_t)ld * i];
    if (d > 0.0) {
      scale[i] = 1.0 / sqrt(d);
    } else {
      scale[i] = 1.0;
      atomicOr(bad, 1);
    }
  }
  double mx = d > 0.0 ? d : 0.0;
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0 && mx > 0.0)
    atomicMax(bmax, (unsigned long long)__double_as_longlong(mx));
}

// A32 = float(S B S) on the lower triangle by 64x64 tiles (grid (nt, nt); the
// tiles above the diagonal return at once), zeros above the diagonal inside
// diagonal tiles
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ B, int ld, int m,
                                                const double* __restrict__ scale,
                                                float* __restrict__ A, int lda) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (J > I) return;
  const int t = threadIdx.x, r = I * 64 + (t & 63);
  const int rr = min(r, m - 1);
  const double sr = scale[rr];
  double v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = B[rr + (size_t)ld * min(J * 64 + (t >> 6) + 4 * k, m - 1)];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = J * 64 + (t >> 6) + 4 * k;
    if (r < m && c < m) A[r + (size_t)lda * c] = r >= c ? (float)(v[k] * sr * scale[c]) : 0.0f;
  }
}

// P[r + ldp c] = A32[r, K0 + c] for rows K0 <= r < m, c < pw (zero above the
// diagonal); grid (ceil((m - K0) / 64), ceil(pw / 64))
__global__ __launch_bounds__(256) void k_panel_to64(const float* __restrict__ A, int lda, int m,
                                                    int K0, int pw, double* __restrict__ P,
                                                    int ldp) {
  const int t = threadIdx.x, r = K0 + blockIdx.x * 64 + (t & 63);
  if (r >= m) return;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = blockIdx.y * 64 + (t >> 6) + 4 * k;
    if (c < pw) P[r + (size_t)ldp * c] = (r >= K0 + c) ? (double)A[r + (size_t)lda * (K0 + c)] : 0.0;
  }
}

// the factored panel back into A32 (lower part), and its 256x256 diagonal
// block (zero above the diagonal) into D (ld 256).  Rows [K0 + 256, K0 + 512)
// of P are then zeroed, element by element after their read: the next
// panel's diagonal block, whose lower part the next-panel update writes
// there (widened) and whose upper part must read as zero.
__global__ __launch_bounds__(256) void k_panel_to32(double* __restrict__ P, int ldp, int m, int K0,
                                                    int pw, float* __restrict__ A, int lda,
                                                    double* __restrict__ D) {
  const int t = threadIdx.x, r = K0 + blockIdx.x * 64 + (t & 63);
  if (r >= m) return;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = blockIdx.y * 64 + (t >> 6) + 4 * k;
    if (c < pw) {
      const bool low = r >= K0 + c;
      const double v = low ? P[r + (size_t)ldp * c] : 0.0;
      if (low) A[r + (size_t)lda * (K0 + c)] = (float)v;
      if (r < K0 + 256) D[(r - K0) + 256 * (size_t)c] = v;
      if (r >= K0 + 256 && r < K0 + 512) P[r + (size_t)ldp * c] = 0.0;
    }
  }
}

// FP32 counterpart of syrk_tile: C[cb + r][cb + c] -= sum_{k < kw}
// A[cb + r][k0 + k] A[cb + c][k0 + k] for one (16 TM) x (16 TN) tile, one wave,
// operands straight from the matrix, D k-steps of loads in flight.  The f32
// MFMA leaves D[i][j] in lane j + 16 (i / 4), register i % 4 (the FP64 MFMA:
// lane j + 16 (i % 4), register i / 4).
template <int TM, int TN, int D>
__device__ __forceinline__ void syrk32_tile(float* __restrict__ Ab, size_t ld, int k0, int kw,
                                            int cb, int rows, int cend, int wm, int wn,
                                            double* __restrict__ W64) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const float* Pp = Ab + cb + ld * k0;
  const float* pr[TM];
  const float* pc[TN];
#pragma unroll
  for (int a = 0; a < TM; ++a) pr[a] = Pp + min(wm + 16 * a + li, rows - 1) + ld * lk;
#pragma unroll
  for (int a = 0; a < TN; ++a) pc[a] = Pp + min(wn + 16 * a + li, cend - 1) + ld * lk;
  v4f acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int c = 0; c < TN; ++c) acc[a][c] = v4f{0.f, 0.f, 0.f, 0.f};
  float ra[D][TM], ca[D][TN];
  const int ns = kw >> 2;
  const size_t st = ld * 4;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const size_t off = st * (d < ns ? d : 0);
#pragma unroll
    for (int a = 0; a < TM; ++a) ra[d][a] = pr[a][off];
#pragma unroll
    for (int a = 0; a < TN; ++a) ca[d][a] = pc[a][off];
  }
  for (int s0 = 0; s0 < ns; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int s = s0 + d;
      if (s < ns) {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int c = 0; c < TN; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[d][c], ra[d][a], acc[a][c], 0, 0, 0);
        if (s + D < ns) {
          const size_t off = st * (s + D);
#pragma unroll
          for (int a = 0; a < TM; ++a) ra[d][a] = pr[a][off];
#pragma unroll
          for (int a = 0; a < TN; ++a) ca[d][a] = pc[a][off];
        }
      }
    }
  }
  float* C = Ab + cb + ld * cb;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int c = 0; c < TN; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gr = wm + 16 * a + li, gc = wn + 16 * c + 4 * lk + v;
        if (gr < rows && gc < cend && gr >= gc) {
          const float nv = C[gr + ld * gc] - acc[a][c][v];
          C[gr + ld * gc] = nv;
          // the next panel, widened for its FP64 factorization (rows absolute,
          // columns from the panel's first, ld = the matrix's)
          if (W64) W64[(cb + gr) + ld * gc] = (double)nv;
        }
      }
}

template <int TM, int TN, int D>
__global__ __launch_bounds__(64) void k_syrk32_update(float* __restrict__ A, int ld, int n, int k0,
                                                      int kw, int cb, int ce,
                                                      const int* __restrict__ info,
                                                      double* __restrict__ W64, int sbc) {
  constexpr int BMt = 16 * TM, BNt = 16 * TN;
  if (info[0] != 0) return;
  const int rows = n - cb, cend = min(ce, n) - cb;
  if (rows <= 0 || cend <= 0) return;
  const int R = (rows + BMt - 1) / BMt, Cn = (cend + BNt - 1) / BNt;
  if (sbc > 0) {  // the XCD-aware blocked order, as in k_syrk_update
    static_assert(BNt % BMt == 0, "column tile a multiple of the row tile");
    const int per = gridDim.x >> 3;
    const int g = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    int i = 0, j = 0;
    if (!blocked_tile(g, R, Cn, BNt / BMt, sbc, i, j)) return;
    syrk32_tile<TM, TN, D>(A, ld, k0, kw, cb, rows, cend, i * BMt, j * BNt, W64);
    return;
  }
  for (int tile = blockIdx.x;; tile += gridDim.x) {
    int g = tile, i = 0, j = 0;
    for (; j < Cn; ++j) {
      const int cnt = R - (j * BNt) / BMt;
      if (g < cnt) { i = (j * BNt) / BMt + g; break; }
      g -= cnt;
    }
    if (j >= Cn) return;
    syrk32_tile<TM, TN, D>(A, ld, k0, kw, cb, rows, cend, i * BMt, j * BNt, W64);
  }
}

template <int TM, int TN, int D>
static void launch_syrk32_t(hipStream_t s, float* A, int ld, int n, int k0, int kw, int cb, int ce,
                            const int* d_info, int maxwaves, double* W64) {
  constexpr int BMt = 16 * TM, BNt = 16 * TN;
  const int rows = n - cb, cols = std::min(ce, n) - cb;
  if (rows <= 0 || cols <= 0) return;
  const int R = (rows + BMt - 1) / BMt, Cn = (cols + BNt - 1) / BNt;
  int ntiles = 0;
  for (int j = 0; j < Cn; ++j) ntiles += R - (j * BNt) / BMt;
  if (maxwaves > 0) ntiles = std::min(ntiles, maxwaves);
  const int sbc = maxwaves <= 0 ? syrk_band() : 0;
  if (sbc) ntiles = (ntiles + 7) / 8 * 8;
  hipLaunchKernelGGL((k_syrk32_update<TM, TN, D>), dim3(ntiles), dim3(64), 0, s, A, ld, n, k0, kw,
                     cb, ce, d_info, W64, sbc);
}

// W64 (optional): also write the updated columns, widened, to the FP64
// panel buffer (the next panel: ce - cb <= 256)
static void launch_syrk32(hipStream_t s, float* A, int ld, int n, int k0, int kw, int cb, int ce,
                          const int* d_info, int maxwaves = 0, double* W64 = nullptr) {
  if (std::min(ce, n) - cb > 256)
    launch_syrk32_t<4, 4, 4>(s, A, ld, n, k0, kw, cb, ce, d_info, maxwaves, nullptr);
  else
    launch_syrk32_t<2, 2, 8>(s, A, ld, n, k0, kw, cb, ce, d_info, maxwaves, W64);
}

int potrf_mixed(hipStream_t s, DenseWork* w, const double* B, int ld, int m, float* A32,
                double* scale, double* P64, double* D64, double* dinv, int* d_info, int* d_bad,
                unsigned long long* d_bmax) {
  constexpr int PW = 256;
  const int nt = (m + 63) / 64, np = (m + PW - 1) / PW;
  SDPA_HIP_CHECK(hipMemsetAsync(d_info, 0, sizeof(int), s));
  SDPA_HIP_CHECK(hipMemsetAsync(d_bad, 0, sizeof(int), s));
  SDPA_HIP_CHECK(hipMemsetAsync(d_bmax, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_eq_scale, dim3((m + 255) / 256), dim3(256), 0, s, B, ld, m, scale, d_bad,
                     d_bmax);
  hipLaunchKernelGGL(k_to_f32, dim3(nt, nt), dim3(256), 0, s, B, ld, m, scale, A32, m);
  // the FP64 panel buffer P64 (m x PW, ld m) is seen as columns [K0, K0 + PW)
  // of an m x m matrix (base P64 - m K0), so the panel kernel indexes it with
  // the global row and column numbers
  // Look-ahead as in potrf_batched: the panel chain (widen, factor, narrow)
  // and the FP32 update of the next panel's columns on `s`, the rest of each
  // trailing update on the CU-masked side stream.  The next panel's update
  // also writes it widened into P64 (free once the narrowing has read it,
  // same stream); it waits for the rest of the previous panel's update,
  // which writes the same columns (evn), and the rest of panel p's update
  // reads the narrowed panel only after it is written (evp).
  if (int e = w->ensure_side(np)) return e;
  SDPA_HIP_CHECK(hipEventRecord(w->evjoin, s));
  SDPA_HIP_CHECK(hipStreamWaitEvent(w->side, w->evjoin, 0));
  for (int p = 0; p < np; ++p) {
    const int K0 = p * PW, pw = std::min(PW, m - K0);
    const dim3 grid((m - K0 + 63) / 64, (pw + 63) / 64);
    if (p == 0) hipLaunchKernelGGL(k_panel_to64, grid, dim3(256), 0, s, A32, m, m, K0, pw, P64, m);
    // the widened panel by the single-launch panel kernel (its columns are
    // already updated: no prologue)
    if (int e = factor_panel_rows(s, w, P64 - (ptrdiff_t)m * K0, m, m, K0, -1, 1, d_info, dinv))
      return e;
    hipLaunchKernelGGL(k_panel_to32, grid, dim3(256), 0, s, P64, m, m, K0, pw, A32, m,
                       D64 + (size_t)p * PW * PW);
    if (K0 + PW >= m) break;
    if (p > 0) SDPA_HIP_CHECK(hipStreamWaitEvent(s, w->evn[p - 1], 0));
    launch_syrk32(s, A32, m, m, K0, PW, K0 + PW, K0 + 2 * PW, d_info, 0, P64);
    SDPA_HIP_CHECK(hipEventRecord(w->evp[p], s));
    SDPA_HIP_CHECK(hipStreamWaitEvent(w->side, w->evp[p], 0));
    if (K0 + 2 * PW < m)
      launch_syrk32(w->side, A32, m, m, K0, PW, K0 + 2 * PW, m, d_info, w->side_waves);
    SDPA_HIP_CHECK(hipEventRecord(w->evn[p], w->side));
  }
  SDPA_HIP_CHECK(hipEventRecord(w->evjoin, w->side));
  SDPA_HIP_CHECK(hipStreamWaitEvent(s, w->evjoin, 0));
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
    // a thread's 32 loads in flight at once, then the LDS stores
    double la[16], wb[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = threadIdx.x + 256 * q, r = e & 63, c = e >> 6;
      la[q] = (r < ib) ? Lm.ptr[i0 + r + (size_t)Lm.ld * (k0 + c)] : 0.0;
      wb[q] = Wm.ptr[k0 + r + (size_t)Wm.ld * (J * TB + c)];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = threadIdx.x + 256 * q, r = e & 63, c = e >> 6;
      As[r + TLD * c] = la[q];
      Bs[r + TLD * c] = wb[q];
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
  {
    double dv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) dv[q] = Dv[threadIdx.x + 256 * q];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = threadIdx.x + 256 * q;
      As[(e & 63) + TLD * (e >> 6)] = dv[q];
    }
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

// Step 1 when the factorization already produced the 64x64 diagonal
// inverses (potrf_batched's dinv, same slot layout): place them on W's
// diagonal and zero the strictly upper blocks of the block row.
// grid = (nblk, nitems).
__global__ __launch_bounds__(256) void k_diag_place(const MatDesc* __restrict__ L,
                                                    const MatDesc* __restrict__ W, int nblk_max,
                                                    const double* __restrict__ dinv) {
  const int b = blockIdx.y, I = blockIdx.x;
  const MatDesc Lm = L[b], Wm = W[b];
  const int k0 = I * TB;
  if (k0 >= Lm.n) return;
  const int kb = min(TB, Lm.n - k0);
  const double* Dv = dinv + ((size_t)b * nblk_max + I) * TB * TB;
  double* Wd = Wm.ptr + k0 + (size_t)Wm.ld * k0;
  double dv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) dv[q] = Dv[threadIdx.x + 256 * q];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q, r = e & 63, c = e >> 6;
    if (r < kb && c < kb) Wd[r + (size_t)Wm.ld * c] = dv[q];
  }
  const int ncol = Wm.n - (k0 + kb);
  for (int e = threadIdx.x; e < kb * ncol; e += 256) {
    const int r = e % kb, c = e / kb;
    Wm.ptr[k0 + r + (size_t)Wm.ld * (k0 + kb + c)] = 0.0;
  }
}

int trtri_batched(hipStream_t s, DenseWork* w, const MatDesc* h_L, const MatDesc* h_W,
                  int nitems, const double* dinv_pre) {
  if (nitems <= 0) return 0;
  int maxn = 0;
  for (int b = 0; b < nitems; ++b) maxn = h_L[b].n > maxn ? h_L[b].n : maxn;
  const int nblk = (maxn + TB - 1) / TB;
  std::vector<MatDesc> lw(h_L, h_L + nitems);
  lw.insert(lw.end(), h_W, h_W + nitems);
  const void* dm = nullptr;
  if (int e = w->desc_copy(s, lw.data(), sizeof(MatDesc) * 2 * nitems, &dm)) return e;
  const MatDesc* d_mats = static_cast<const MatDesc*>(dm);
  const double* dv = dinv_pre;
  if (dv) {
    hipLaunchKernelGGL(k_diag_place, dim3(nblk, nitems), dim3(256), 0, s, d_mats,
                       d_mats + nitems, nblk, dv);
  } else {
    if (int e = w->ensure_dinv((size_t)nitems * nblk * TB * TB)) return e;
    hipLaunchKernelGGL(k_diag_inv, dim3(nblk, nitems), dim3(64), 0, s, d_mats,
                       d_mats + nitems, nblk, w->d_dinv);
    dv = w->d_dinv;
  }
  for (int I = 1; I < nblk; ++I)
    hipLaunchKernelGGL(k_trtri_row, dim3(I, nitems), dim3(256), 0, s, d_mats,
                       d_mats + nitems, I, nblk, dv);
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
                        double* Winv, const double* dinv64) {
  const int nb = (m + SB - 1) / SB;
  std::vector<MatDesc> Ls(nb), Ws(nb);
  for (int b = 0; b < nb; ++b) {
    const int k0 = b * SB;
    Ls[b] = MatDesc{const_cast<double*>(L) + k0 + (size_t)ld * k0, std::min(SB, m - k0), ld};
    Ws[b] = MatDesc{Winv + (size_t)b * SB * SB, std::min(SB, m - k0), SB};
  }
  // the 64x64 inverses of the factorization (slot k0/64) are slot
  // (256-block b, 64-block I) = 4 b + I of the 256-block batch
  return trtri_batched(s, w, Ls.data(), Ws.data(), nb, dinv64);
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

// ---------------------------------------------------------------------------
// The same two solves, each as ONE cooperative launch (all work-groups
// co-resident): work-group s owns the 64-row strip s; the 256-row block b is
// strips 4b..4b+3.  Forward (L t = x): a strip subtracts L[strip, block b']
// t_b' for every earlier block as soon as t_b' is published, writes its rows
// of y_b = x_b - sum, and once the block's four strips have published y_b it
// forms its rows of t_b = W_b y_b and publishes them.  Backward (L^T x = t)
// runs the blocks from the last to the first with the columns of L.  Blocks
// are published by release stores of a per-strip flag (value = this launch's
// epoch, so flags never need clearing); readers poll with relaxed loads and fence once (acquire),
// then read the published vector with device-scope atomic loads.  Every poll
// is bounded: after kSpinLimit polls the work-group records the failure in
// *err and returns, so a launch always ends.
constexpr unsigned kSpinLimit = 1u << 21;  // about a second of polling

struct TrsvParams {
  const double* L;
  int ld, m;
  const double* W;    // 256 x 256 diagonal-block inverses, ld 256
  const double* rhs;  // forward: x; backward: t
  double* y;          // m-vector scratch (per-block right-hand sides)
  double* out;        // forward: t; backward: x
  int* flags;         // [0, ns): y published, [ns, 2 ns): out published
  int epoch;
  int* err;
};

// thread 0 polls flags[lo, hi) until each equals epoch; false if it gave up
// (or another work-group already had).  The polls are relaxed loads and one
// acquire fence follows the last: an acquire load per poll invalidates the
// L2 each time, which made a hop ~4.7 us instead of ~1.4 us with 94 pollers.
__device__ bool trsv_wait(const int* flags, int lo, int hi, int epoch, int* err) {
  __shared__ int ok;
  if (threadIdx.x == 0) {
    int good = 1;
    for (int f = lo; f < hi && good; ++f) {
      unsigned spins = 0;
      while (__hip_atomic_load(&flags[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > kSpinLimit ||
            ((spins & 63) == 0 &&
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          good = 0;
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

__device__ void trsv_publish(int* flag, int epoch) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// v[0..n) <- src[0..n) through device-scope loads (n <= 256, one per thread)
__device__ void trsv_fetch(double* v, const double* src, int n) {
  const int t = threadIdx.x;
  v[t] = (t < n) ? __hip_atomic_load(const_cast<double*>(src) + t, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT)
                 : 0.0;
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_trsv_fwd(TrsvParams P) {
  const int ns = gridDim.x, s = blockIdx.x, b = s >> 2, q = s & 3;
  const int k0 = b * SB, kb = min(SB, P.m - k0);
  const int t = threadIdx.x, i = t & 63, w = t >> 6;
  const int r = s * 64 + i;
  __shared__ double part[4][64];
  __shared__ double vb[SB];
  double acc = 0.0;
  // rows of the strip times the earlier blocks' solutions
  for (int bp = 0; bp < b; ++bp) {
    // the matrix entries do not depend on the flag: load them first, so their
    // latency overlaps the wait (rows past m read a clamped, valid row and are
    // never stored)
    const double* Lr = P.L + min(r, P.m - 1) + (size_t)P.ld * (bp * SB + 64 * w);
    double lv[64];
#pragma unroll
    for (int c = 0; c < 64; ++c) lv[c] = Lr[(size_t)P.ld * c];
    if (!trsv_wait(P.flags + ns, 4 * bp, 4 * bp + 4, P.epoch, P.err)) return;
    trsv_fetch(vb, P.out + bp * SB, SB);
    const double* v = vb + 64 * w;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 64; ++c) a[c & 3] += lv[c] * v[c];
    acc += (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();  // vb is refilled next round
  }
  part[w][i] = acc;
  __syncthreads();
  if (w == 0 && r < P.m)
    P.y[r] = P.rhs[r] - ((part[0][i] + part[1][i]) + (part[2][i] + part[3][i]));
  trsv_publish(&P.flags[s], P.epoch);
  // rows 64 q .. 64 q + 63 of t_b = W_b y_b (W_b lower triangular; only the
  // column quarters w <= q are nonzero).  Columns past kb: a clamped (valid)
  // entry of W times vb = 0; rows past kb read a clamped row, never stored.
  const int row = 64 * q + i;
  double lw[64];
  if (w <= q) {
    const double* Wr = P.W + (size_t)b * SB * SB + min(row, kb - 1);
#pragma unroll
    for (int c = 0; c < 64; ++c) lw[c] = Wr[(size_t)SB * min(64 * w + c, kb - 1)];
  }
  if (!trsv_wait(P.flags, 4 * b, min(4 * b + 4, ns), P.epoch, P.err)) return;
  trsv_fetch(vb, P.y + k0, kb);
  acc = 0.0;
  if (w <= q) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 64; ++c) a[c & 3] += lw[c] * vb[64 * w + c];
    acc = (a[0] + a[1]) + (a[2] + a[3]);
  }
  part[w][i] = acc;
  __syncthreads();
  if (w == 0 && r < P.m) P.out[r] = (part[0][i] + part[1][i]) + (part[2][i] + part[3][i]);
  trsv_publish(&P.flags[ns + s], P.epoch);
}

// backward: columns of L (contiguous); wave w owns output columns
// 16 w .. 16 w + 15 of the strip, lanes run down the rows of a block and the
// 16 per-lane partial sums are reduced through LDS at the end
__global__ __launch_bounds__(256) void k_trsv_bwd(TrsvParams P) {
  const int ns = gridDim.x, s = blockIdx.x, b = s >> 2, q = s & 3;
  const int nb = (P.m + SB - 1) / SB;
  const int k0 = b * SB, kb = min(SB, P.m - k0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int c0 = s * 64 + 16 * w;  // first output column of this wave
  __shared__ double red[4][16][65];
  __shared__ double vb[SB];
  __shared__ double col[64];
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  for (int bp = nb - 1; bp > b; --bp) {
    const int kbp = min(SB, P.m - bp * SB);
    // the block's 64 matrix entries per thread are loaded before the wait;
    // rows past the block and columns past m read clamped (valid) entries:
    // the former are multiplied by vb = 0, the latter are never stored
    double lv[4][16];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const double* Lc = P.L + (bp * SB + min(64 * rr + lane, kbp - 1));
#pragma unroll
      for (int j = 0; j < 16; ++j) lv[rr][j] = Lc[(size_t)P.ld * min(c0 + j, P.m - 1)];
    }
    if (!trsv_wait(P.flags + ns, 4 * bp, min(4 * bp + 4, ns), P.epoch, P.err)) return;
    trsv_fetch(vb, P.out + bp * SB, kbp);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const double xv = vb[64 * rr + lane];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] += lv[rr][j] * xv;
    }
    __syncthreads();
  }
  // reduce the 16 columns over the lanes
#pragma unroll
  for (int j = 0; j < 16; ++j) red[w][j][lane] = acc[j];
  __syncthreads();
  {
    const int j = t >> 2, part4 = t & 3;  // column j of the strip, quarter of the lanes
    const int ww = j >> 4, jj = j & 15;
    double sacc = 0.0;
    for (int l = 16 * part4; l < 16 * part4 + 16; ++l) sacc += red[ww][jj][l];
    sacc += __shfl_xor(sacc, 1, 64);
    sacc += __shfl_xor(sacc, 2, 64);
    const int r = s * 64 + j;
    if (part4 == 0 && r < P.m) P.y[r] = P.rhs[r] - sacc;
  }
  trsv_publish(&P.flags[s], P.epoch);
  // rows 64 q + j of x_b = W_b^T y_b: column 64 q + j of W_b, rows c >= 64 q + j
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  {
    double lv[4][16];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const double* Wc = P.W + (size_t)b * SB * SB + min(64 * rr + lane, kb - 1);
#pragma unroll
      for (int j = 0; j < 16; ++j) lv[rr][j] = Wc[(size_t)SB * min(64 * q + 16 * w + j, kb - 1)];
    }
    if (!trsv_wait(P.flags, 4 * b, min(4 * b + 4, ns), P.epoch, P.err)) return;
    trsv_fetch(vb, P.y + k0, kb);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const double yv = vb[64 * rr + lane];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] += lv[rr][j] * yv;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) red[w][j][lane] = acc[j];
  __syncthreads();
  {
    const int j = t >> 2, part4 = t & 3;
    const int ww = j >> 4, jj = j & 15;
    double sacc = 0.0;
    for (int l = 16 * part4; l < 16 * part4 + 16; ++l) sacc += red[ww][jj][l];
    sacc += __shfl_xor(sacc, 1, 64);
    sacc += __shfl_xor(sacc, 2, 64);
    const int r = s * 64 + j;
    if (part4 == 0 && r < P.m) col[j] = sacc;
    __syncthreads();
    if (t < 64 && s * 64 + t < P.m) P.out[s * 64 + t] = col[t];
  }
  trsv_publish(&P.flags[ns + s], P.epoch);
}

// ---------------------------------------------------------------------------
// Pipelined solves.  The same blocks (256 rows), strips (64 rows, one
// work-group each) and diagonal-block inverses W_b as above, with three
// changes that shorten each step of the chain:
//  - values pass between work-groups as epoch-tagged 64-bit words: value v
//    of a launch with epoch e is the pair (lo32(v) | e << 32, hi32(v) | e << 32),
//    each word written by one atomic store, so a reader that sees e in both
//    words has the value: no flag, fence or second read;
//  - one wave of each work-group (the poller) does all the waiting and
//    fetching, into an LDS copy of every finished block of the solution,
//    while eight compute waves run the products;
//  - the compute waves issue the loads of the next tile of L before they use
//    the current one, so a strip's reads of L stream while the chain advances.
// 576 threads: compute waves 0..7, poller wave 8.  Every wave passes the
// same barriers: the loops run on a count the poller posts in LDS before
// each barrier.
constexpr int kLLThreads = 576;


// Raw buffer loads: a 32-bit per-lane byte offset plus a wave-uniform one in a
// scalar register, so 32 loads in flight cost 32 data registers and one
// offset register rather than a 64-bit address each.  (Buffers of < 2 GiB;
// reads past num_bytes return 0.)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, size_t num_bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)num_bytes,
                                           0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// an element of a factor stored as T (double, or float for the mixed-precision
// factor), widened to double
template <typename T>
__device__ __forceinline__ double buf_ldT(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ double buf_ldT<double>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                  unsigned soff) {
  return buf_ld(r, voff, soff);
}
template <>
__device__ __forceinline__ double buf_ldT<float>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                 unsigned soff) {
  return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

struct LLParams {
  const void* L;              // the factor, double or float (kernel template)
  int ld, m;
  const double* W;            // 256 x 256 diagonal-block inverses, ld 256
  const double* rhs;          // forward: x; backward: t (plain vectors)
  double* out;                // forward: t; backward: x
  unsigned long long* ybuf;   // tagged words of y (2 per value)
  unsigned long long* sbuf;   // tagged words of the block solutions
  unsigned epoch;
  int* err;
};

__device__ __forceinline__ void ll_put(unsigned long long* p, double v, unsigned e) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned long long tag = (unsigned long long)e << 32;
  __hip_atomic_store(p, (u & 0xffffffffull) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, (u >> 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// poller wave: values [i0, i0 + 256) of a tagged buffer into dst (LDS),
// indices >= m as 0.  Waits until all carry epoch e; false if it gave up
// (spin limit, or another work-group already had).  It polls one word per
// 64-value strip (the strip's last value) and reads the block only once all
// four show e, so that the many work-groups waiting on one block do not
// flood its memory channel; dist (blocks between this one and the one the
// caller needs next) lengthens the pause between polls of blocks that are
// not yet urgent.  The block needed next (dist 0) is read whole each time.
__device__ bool ll_get_block(const unsigned long long* buf, int i0, int m, unsigned e, double* dst,
                             int lane, int* err, int dist) {
  unsigned spins = 0;
  const size_t probe = (size_t)min(i0 + 64 * (lane & 3) + 63, m - 1);
  while (true) {
    // the block needed next is read whole at once (one round trip); the
    // others are probed first
    const unsigned long long pw =
        dist == 0 ? 0ull
                  : __hip_atomic_load(buf + 2 * probe + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (dist == 0 || __all((unsigned)(pw >> 32) == e)) {
      unsigned long long w0[4], w1[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t idx = (size_t)min(i0 + lane + 64 * k, m - 1);
        w0[k] = __hip_atomic_load(buf + 2 * idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w1[k] = __hip_atomic_load(buf + 2 * idx + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // written straight to dst: if some word is stale the block is read again
      // (nobody reads dst before this returns true)
      bool ok = true;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ok = ok && (unsigned)(w0[k] >> 32) == e && (unsigned)(w1[k] >> 32) == e;
        const long long bits = (long long)((w0[k] & 0xffffffffull) | (w1[k] << 32));
        dst[lane + 64 * k] = (i0 + lane + 64 * k < m) ? __longlong_as_double(bits) : 0.0;
      }
      if (__all(ok)) break;
    }
    ++spins;
    if (spins > kSpinLimit ||
        ((spins & 63) == 0 &&
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
    for (int d = 0; d < dist; ++d) __builtin_amdgcn_s_sleep(8);
  }
  return true;
}

// forward, L t = x.  Strip s (rows 64 s + lane) of block b: y = x - sum over
// earlier blocks of L[rows, block] t_block (compute wave w takes columns
// 32 w .. 32 w + 31 of each block), then t_b = W_b y_b for its rows.
template <typename T>
__global__ __launch_bounds__(kLLThreads) void k_trsv_fwd_ll(LLParams P) {
  extern __shared__ double sol[];  // finished blocks of t
  __shared__ double ysh[SB];
  __shared__ double part[8][64];
  __shared__ int post[2];
  __shared__ int yok;
  const int s = blockIdx.x, b = s >> 2, q = s & 3;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const bool poller = w == 8;
  const int m = P.m, k0 = b * SB, kb = min(SB, m - k0);
  const int r = s * 64 + lane;
  // wave-uniform column bases (scalar registers) plus a 32-bit per-lane row
  // offset: rows past m read a valid row and are never stored
  const int wc = min(w, 7);
  const unsigned rl = (unsigned)sizeof(T) * (unsigned)min(r, m - 1);
  const unsigned ldb = (unsigned)sizeof(T) * (unsigned)P.ld;
  const __amdgpu_buffer_rsrc_t Lrs = buf_rsrc(P.L, (size_t)P.ld * m * sizeof(T));
  // this thread's W_b entries: row 64 q + lane, columns 32 w + c, clamped into
  // the kb x kb block (clamped columns meet ysh = 0, clamped rows are not stored)
  const __amdgpu_buffer_rsrc_t Wrs = buf_rsrc(P.W + (size_t)b * SB * SB, SB * SB * sizeof(double));
  const unsigned wl = 8u * (unsigned)min(64 * q + lane, kb - 1);
  double lv[32];
  double acc = 0.0;
#define LL_LOAD_L(dst, bp)                                                   \
  _Pragma("unroll") for (int c = 0; c < 32; ++c)                             \
    dst[c] = buf_ldT<T>(Lrs, rl, ldb * (unsigned)((bp) * SB + 32 * wc + c));
#define LL_LOAD_W(dst)                                                       \
  _Pragma("unroll") for (int c = 0; c < 32; ++c)                             \
    dst[c] = buf_ld(Wrs, wl, 8u * SB * (unsigned)min(32 * wc + c, kb - 1));
  // lw: this thread's W_b tile (waves whose columns meet the lower triangle),
  // loaded once up front.  lv: the background tile of the next round.  A
  // round multiplies lv and then issues the loads of the following tile into
  // lv, so they land while the work-group waits for the next block (a second
  // buffer would need register copies, whose reads wait for the loads just
  // issued).  The last tile, the one the chain waits for, is multiplied after
  // the loop with nothing issued behind it: 128 KB of loads take the CU about
  // a microsecond to issue.
  double lw[32];
  const double xr = (w == 0 && r < m) ? P.rhs[r] : 0.0;  // read early: off the chain
  if (!poller) {
    if (w <= 2 * q + 1) LL_LOAD_W(lw)
    if (b > 0) LL_LOAD_L(lv, 0)
  }
  // lv times 32 values of the LDS copy, read 8 at a time ahead of the products
#define LL_FMA_TILE(tvp)                                                      \
  {                                                                           \
    const double* tv_ = (tvp);                                                \
    double a[4] = {0.0, 0.0, 0.0, 0.0};                                       \
    _Pragma("unroll") for (int h8 = 0; h8 < 32; h8 += 8) {                    \
      double t8[8];                                                           \
      _Pragma("unroll") for (int c = 0; c < 8; ++c) t8[c] = tv_[h8 + c];      \
      _Pragma("unroll") for (int c = 0; c < 8; ++c)                           \
        a[c & 3] = fma(lv[h8 + c], t8[c], a[c & 3]);                          \
    }                                                                         \
    acc += (a[0] + a[1]) + (a[2] + a[3]);                                     \
  }
  int have = 0, done = 0;  // tiles fetched (poller) / multiplied (compute waves)
  for (int it = 0;; ++it) {
    if (poller) {
      int h = have;
      if (h < b)
        h = ll_get_block(P.sbuf, h * SB, m, P.epoch, sol + h * SB, lane, P.err, min(b - 1 - h, 8))
                ? h + 1
                : -1;
      if (lane == 0) post[it & 1] = h;
    }
    lds_barrier();
    const int h = __builtin_amdgcn_readfirstlane(post[it & 1]);  // uniform: scalar offsets
    if (h < 0) return;
    if (!poller) {
      for (; done < h && done < b - 1; ++done) {
        LL_FMA_TILE(sol + done * SB + 32 * wc)
        // keep the reloads below the products (hoisted, they would need a
        // second set of registers); the reload is of a tile <= b - 1
        __asm__ volatile("" ::: "memory");
        LL_LOAD_L(lv, done + 1)
      }
    }
    have = h;
    if (h >= b) break;
  }
  if (!poller && b > 0) LL_FMA_TILE(sol + (b - 1) * SB + 32 * wc)
#undef LL_FMA_TILE
#undef LL_LOAD_L
#undef LL_LOAD_W
  // y for the strip's rows
  if (!poller) part[w][lane] = acc;
  lds_barrier();
  if (w == 0 && r < m) {
    const double sum = ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) +
                       ((part[4][lane] + part[5][lane]) + (part[6][lane] + part[7][lane]));
    ll_put(P.ybuf + 2 * (size_t)r, xr - sum, P.epoch);
  }
  if (poller) {
    const bool ok = ll_get_block(P.ybuf, k0, m, P.epoch, ysh, lane, P.err, 0);
    if (lane == 0) yok = ok ? 1 : 0;
  }
  lds_barrier();
  if (!yok) return;
  // t_b rows: W_b[64 q + lane, :] y_b (columns past 64 q + 63 are zero)
  double u = 0.0;
  if (!poller && w <= 2 * q + 1) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 32; ++c) a[c & 3] = fma(lw[c], ysh[32 * w + c], a[c & 3]);
    u = (a[0] + a[1]) + (a[2] + a[3]);
  }
  if (!poller) part[w][lane] = u;
  lds_barrier();
  if (w == 0 && r < m) {
    const double tv = ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) +
                      ((part[4][lane] + part[5][lane]) + (part[6][lane] + part[7][lane]));
    ll_put(P.sbuf + 2 * (size_t)r, tv, P.epoch);
    P.out[r] = tv;
  }
}

// backward, L^T x = t.  Strip s (columns 64 s .. 64 s + 63 of L) of block b:
// y = t - sum over later blocks of L[block, columns]^T x_block (compute wave
// w takes columns 64 s + 8 w + j, lanes run down the block's rows), then
// x_b = W_b^T y_b for its columns.
__device__ __forceinline__ double ll_colsum(double (*red)[8][65], int tid) {
  // thread tid < 512: column cc = tid >> 3 of the strip, lanes 8 p .. 8 p + 7
  const int cc = tid >> 3, p = tid & 7;
  const double* rr = red[cc >> 3][cc & 7] + 8 * p;
  double v = ((rr[0] + rr[1]) + (rr[2] + rr[3])) + ((rr[4] + rr[5]) + (rr[6] + rr[7]));
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(kLLThreads) void k_trsv_bwd_ll(LLParams P) {
  extern __shared__ double sol[];  // finished blocks of x
  __shared__ double ysh[SB];
  __shared__ double red[8][8][65];
  __shared__ int post[2];
  __shared__ int yok;
  const int s = blockIdx.x, b = s >> 2, q = s & 3;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const bool poller = w == 8;
  const int m = P.m, nb = (m + SB - 1) / SB, k0 = b * SB, kb = min(SB, m - k0);
  const int wc = min(w, 7);
  const int c0 = s * 64 + 8 * wc;  // this wave's first column
  const int nbg = nb - 1 - b;      // background tiles: block nb - 1 - i, i < nbg
  // columns c0 + j are < m whenever there is background (b < nb - 1); the
  // column bases are wave-uniform, the rows a 32-bit per-lane offset
  const unsigned ldb = (unsigned)sizeof(T) * (unsigned)P.ld;
  const unsigned cb = ldb * (unsigned)min(c0, m - 8);
  const __amdgpu_buffer_rsrc_t Lrs = buf_rsrc(P.L, (size_t)P.ld * m * sizeof(T));
  const __amdgpu_buffer_rsrc_t Wrs = buf_rsrc(P.W + (size_t)b * SB * SB, SB * SB * sizeof(double));
  double lv[32];
  double acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0;
  // W_b^T tile: rows 64 k + lane of W_b (k >= q: the rows above are zero),
  // columns 64 q + 8 w + j, clamped into the kb x kb block (clamped rows meet
  // ysh = 0, clamped columns are not stored)
#define LL_LOAD_W(dst)                                                                   \
  _Pragma("unroll") for (int k = 0; k < 4; ++k)                                          \
    _Pragma("unroll") for (int j = 0; j < 8; ++j)                                        \
      dst[8 * k + j] = (k >= q) ? buf_ld(Wrs, 8u * (unsigned)min(64 * k + lane, kb - 1),  \
                                         8u * SB * (unsigned)min(64 * q + 8 * wc + j, kb - 1)) \
                                : 0.0;
#define LL_LOAD_L(dst, bp)                                                               \
  _Pragma("unroll") for (int k = 0; k < 4; ++k)                                          \
    _Pragma("unroll") for (int j = 0; j < 8; ++j)                                        \
      dst[8 * k + j] = buf_ldT<T>(Lrs, (unsigned)sizeof(T) * (unsigned)min((bp) * SB + 64 * k + lane, m - 1), \
                                  cb + ldb * j);
  // lv: the next background tile, reloaded right after each use (as in the
  // forward kernel).  The W_b^T tile (lw) is loaded once y_b is published,
  // so the loads overlap its exchange: held through the loop it would not
  // fit in registers beside lv, and issued before the last tile's products
  // it would delay the chain.
  double lw[32];
  const int col = s * 64 + (tid >> 3);  // the column ll_colsum gives this thread
  const double xr = (!poller && (tid & 7) == 0 && col < m) ? P.rhs[col] : 0.0;  // read early
  if (!poller && nbg > 0) LL_LOAD_L(lv, nb - 1)
  int have = 0, done = 0;  // tiles fetched (poller) / multiplied (compute waves)
  for (int it = 0;; ++it) {
    if (poller) {
      int h = have;
      if (h < nbg) {
        const int bp = nb - 1 - h;
        h = ll_get_block(P.sbuf, bp * SB, m, P.epoch, sol + bp * SB, lane, P.err,
                         min(bp - b - 1, 8))
                ? h + 1
                : -1;
      }
      if (lane == 0) post[it & 1] = h;
    }
    lds_barrier();
    const int h = __builtin_amdgcn_readfirstlane(post[it & 1]);  // uniform: scalar offsets
    if (h < 0) return;
    if (!poller) {
      // all but the last tile here; every reload is of a tile < nbg
      for (; done < h && done < nbg - 1; ++done) {
        const double* xv = sol + (nb - 1 - done) * SB + lane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double xk = xv[64 * k];
#pragma unroll
          for (int j = 0; j < 8; ++j) acc[j] = fma(lv[8 * k + j], xk, acc[j]);
        }
        __asm__ volatile("" ::: "memory");  // reloads stay below the products
        LL_LOAD_L(lv, nb - 2 - done)
      }
    }
    have = h;
    if (h >= nbg) break;
  }
  if (!poller) {
    if (nbg > 0) {  // the last tile: block b + 1
      const double* xv = sol + (b + 1) * SB + lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double xk = xv[64 * k];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = fma(lv[8 * k + j], xk, acc[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[w][j][lane] = acc[j];
  }
  lds_barrier();
  if (!poller) {
    const double v = ll_colsum(red, tid);
    if ((tid & 7) == 0 && col < m) ll_put(P.ybuf + 2 * (size_t)col, xr - v, P.epoch);
    LL_LOAD_W(lw)
  } else {
    const bool ok = ll_get_block(P.ybuf, k0, m, P.epoch, ysh, lane, P.err, 0);
    if (lane == 0) yok = ok ? 1 : 0;
  }
  lds_barrier();
  if (!yok) return;
#undef LL_LOAD_W
#undef LL_LOAD_L
  if (!poller) {
    double a2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a2[j] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double yk = ysh[64 * k + lane];
#pragma unroll
      for (int j = 0; j < 8; ++j) a2[j] = fma(lw[8 * k + j], yk, a2[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[w][j][lane] = a2[j];
  }
  lds_barrier();
  if (!poller) {
    const double v = ll_colsum(red, tid);
    if ((tid & 7) == 0 && col < m) {
      ll_put(P.sbuf + 2 * (size_t)col, v, P.epoch);
      P.out[col] = v;
    }
  }
}

// x := L^-T L^-1 x by the two pipelined cooperative launches; returns 1 when
// they do not apply (no cooperative launch, or the grid cannot be co-resident
// with its LDS) and the caller falls back
template <typename T>
static int chol_solve_ll(hipStream_t s, DenseWork* w, const T* L, int ld, int m,
                         const double* Winv, double* x, const double* rhs = nullptr) {
  const int ns = (m + 63) / 64, nb = (m + SB - 1) / SB;
  const size_t lds = (size_t)nb * SB * sizeof(double);
  if (nb > 48 || (size_t)ld * m * sizeof(T) >= (size_t)1 << 31) return 1;
  int& max_nb = sizeof(T) == 8 ? w->ll_max_nb : w->ll_max_nb32;
  if (max_nb < nb) {
    int dev = 0, coop = 0, ncu = 0;
    SDPA_HIP_CHECK(hipGetDevice(&dev));
    SDPA_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    SDPA_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (!coop) return 1;
    SDPA_HIP_CHECK(hipFuncSetAttribute((const void*)k_trsv_fwd_ll<T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SDPA_HIP_CHECK(hipFuncSetAttribute((const void*)k_trsv_bwd_ll<T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int pf = 0, pb = 0;
    SDPA_HIP_CHECK(
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&pf, k_trsv_fwd_ll<T>, kLLThreads, lds));
    SDPA_HIP_CHECK(
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&pb, k_trsv_bwd_ll<T>, kLLThreads, lds));
    if ((long)std::min(pf, pb) * ncu < ns) return 1;
    max_nb = nb;
  }
  if (int e = w->ensure_vec((size_t)m)) return e;
  if (int e = w->ensure_flags(0)) return e;  // the failure flag
  if (int e = w->ensure_ll(4 * (size_t)nb * SB)) return e;
  LLParams P;
  P.L = L; P.ld = ld; P.m = m; P.W = Winv; P.err = w->d_err;
  P.ybuf = w->d_ll;
  P.sbuf = w->d_ll + 2 * (size_t)nb * SB;
  double* t = w->d_vec;
  // forward: t = L^-1 x
  if (++w->ll_epoch == 0) ++w->ll_epoch;
  P.rhs = rhs ? rhs : x; P.out = t; P.epoch = w->ll_epoch;
  void* args[] = {&P};
  SDPA_HIP_CHECK(hipLaunchCooperativeKernel((const void*)k_trsv_fwd_ll<T>, dim3(ns), dim3(kLLThreads),
                                            args, lds, s));
  // backward: x = L^-T t
  LLParams Q = P;
  if (++w->ll_epoch == 0) ++w->ll_epoch;
  Q.rhs = t; Q.out = x; Q.epoch = w->ll_epoch;
  void* args2[] = {&Q};
  SDPA_HIP_CHECK(hipLaunchCooperativeKernel((const void*)k_trsv_bwd_ll<T>, dim3(ns), dim3(kLLThreads),
                                            args2, lds, s));
  SDPA_HIP_CHECK(hipMemcpyAsync(w->h_err, w->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  return 0;
}

// x := L^-T L^-1 x by the two cooperative launches; returns 1 when the
// device cannot launch them cooperatively (the caller falls back)
static int chol_solve_coop(hipStream_t s, DenseWork* w, const double* L, int ld, int m,
                           const double* Winv, double* x, const double* rhs = nullptr) {
  if (w->coop_state == 0) {
    int dev = 0, coop = 0;
    SDPA_HIP_CHECK(hipGetDevice(&dev));
    SDPA_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    w->coop_state = coop ? 1 : 2;
  }
  if (w->coop_state != 1) return 1;
  const int ns = (m + 63) / 64;
  if (int e = w->ensure_vec(3 * (size_t)m)) return e;
  if (int e = w->ensure_flags(2 * ns)) return e;
  TrsvParams P;
  P.L = L; P.ld = ld; P.m = m; P.W = Winv; P.flags = w->d_flags; P.err = w->d_err;
  double* t = w->d_vec;
  double* y = w->d_vec + m;
  // forward: t = L^-1 x
  P.rhs = rhs ? rhs : x; P.y = y; P.out = t; P.epoch = ++w->epoch;
  void* args[] = {&P};
  SDPA_HIP_CHECK(hipLaunchCooperativeKernel((const void*)k_trsv_fwd, dim3(ns), dim3(256), args, 0, s));
  // backward: x = L^-T t
  TrsvParams Q = P;
  Q.rhs = t; Q.out = x; Q.epoch = ++w->epoch;
  void* args2[] = {&Q};
  SDPA_HIP_CHECK(hipLaunchCooperativeKernel((const void*)k_trsv_bwd, dim3(ns), dim3(256), args2, 0, s));
  SDPA_HIP_CHECK(hipMemcpyAsync(w->h_err, w->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  return 0;
}

int chol_solve(hipStream_t s, DenseWork* w, const double* L, int ld, int m, const double* Winv,
               double* x, const double* rhs) {
  if (m <= 0) return 0;
  if (rhs == x) rhs = nullptr;
  if (m > 2 * SB) {
    int r = chol_solve_ll<double>(s, w, L, ld, m, Winv, x, rhs);
    if (r <= 0) return r;
    r = chol_solve_coop(s, w, L, ld, m, Winv, x, rhs);
    if (r <= 0) return r;
  }
  if (int e = w->ensure_vec((size_t)m)) return e;
  double* t = w->d_vec;
  const int nb = (m + SB - 1) / SB;
  // one block: the forward step only reads the right-hand side; more: the
  // forward updates consume it in x
  if (rhs && nb > 1) {
    SDPA_HIP_CHECK(hipMemcpyAsync(x, rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    rhs = nullptr;
  }
  // forward: t = L^-1 x   (x is consumed as the right-hand side)
  for (int b = 0; b < nb; ++b) {
    hipLaunchKernelGGL(k_block_gemv, dim3(8), dim3(256), 0, s, Winv, m, b, 0, rhs ? rhs : x, t);
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

// Whether chol_solve_f32 applies to an m x m float factor with leading
// dimension ld (the pipelined kernels' size range, cooperative launch).
bool chol_solve_f32_supported(int m, int ld) {
  const int nb = (m + SB - 1) / SB;
  if (m <= 2 * SB || nb > 48 || (size_t)ld * m * sizeof(float) >= (size_t)1 << 31) return false;
  int dev = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess)
    return false;
  return coop != 0;
}

// x := L^-T L^-1 x with a float factor (the mixed-precision path); the
// pipelined cooperative kernels only.  Returns 1 when they do not apply.
int chol_solve_f32(hipStream_t s, DenseWork* w, const float* L, int ld, int m, const double* Winv,
                   double* x) {
  if (m <= 0) return 0;
  if (m <= 2 * SB) return 1;
  return chol_solve_ll<float>(s, w, L, ld, m, Winv, x);
}

// ---------------------------------------------------------------------------
// Tile-sparse factorization of the sparse-Schur branch (the tables come from
// sparse_schur_analyse).  The permuted, shifted Schur matrix is gathered into
// the stored tiles of its factor's pattern (64 x 64, column-major; rows and
// columns beyond m are the identity); then, level by level of the tile
// elimination tree: every diagonal tile of the level is factored with its
// inverse (the 64 x 64 routines of the dense factorization, LAPACK's pivot
// rule), every off-diagonal tile of the level's columns solved,
// L(I, J) = A(I, J) inv(L_JJ)^T, and every target tile of the level's
// updates gets A(I, K) -= sum_J L(I, J) L(K, J)^T from one work-group, its
// sources in column order.

__global__ __launch_bounds__(256) void k_ts_gather(const double* __restrict__ B, int m,
                                                   const int* __restrict__ perm,
                                                   const int* __restrict__ tI,
                                                   const int* __restrict__ tJ,
                                                   double* __restrict__ tiles, double shift) {
  const int p = blockIdx.x;
  const int I = tI[p], J = tJ[p];
  double* D = tiles + (size_t)p * TB * TB;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = 64 * I + (e & 63), c = 64 * J + (e >> 6);
    double v;
    if (r < m && c < m) {
      const int pr = perm[r], pc = perm[c];
      v = B[(size_t)max(pr, pc) + (size_t)m * min(pr, pc)] + (r == c ? shift : 0.0);
    } else {
      v = (r == c) ? 1.0 : 0.0;
    }
    D[e] = v;
  }
}

// the diagonal tiles of one level: factor (LAPACK's rule) and invert
__global__ __launch_bounds__(256) void k_ts_potrf(double* __restrict__ tiles,
                                                  const int* __restrict__ colptr,
                                                  const int* __restrict__ cols,
                                                  double* __restrict__ winv, int* __restrict__ info) {
  const int J = cols[blockIdx.x];
  const MatDesc M{tiles + (size_t)colptr[J] * TB * TB, TB, TB};
  diag_block<0>(M, 0, 0, -1, info + J, winv + (size_t)J * TB * TB, 1);
}

// the off-diagonal tiles of one level's columns: L(I, J) = A(I, J) W_J^T, in
// place; wave w takes rows 16 w .. 16 w + 15 (each wave reads only the rows
// it writes)
__global__ __launch_bounds__(256) void k_ts_trsm(double* __restrict__ tiles,
                                                 const int* __restrict__ list,
                                                 const int* __restrict__ tJ,
                                                 const double* __restrict__ winv) {
  const int p = list[blockIdx.x];
  double* X = tiles + (size_t)p * TB * TB;
  const double* Wt = winv + (size_t)tJ[p] * TB * TB;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int row = 16 * w + li;
  double a[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) a[s] = X[row + TB * (4 * s + lk)];
  v4d acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 16; ++s)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(Wt[(16 * c + li) + TB * (4 * s + lk)], a[s],
                                                    acc[c], 0, 0, 0);
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) X[row + TB * (16 * c + 4 * v + lk)] = acc[c][v];
}

// the update tasks of one level: target tile -= sum over its source pairs of
// L(I, J) L(K, J)^T, or (slot >= 0) that sum stored to the partial slot
__global__ __launch_bounds__(256) void k_ts_update(double* __restrict__ tiles,
                                                   const int* __restrict__ target,
                                                   const int* __restrict__ src_ptr,
                                                   const int* __restrict__ src,
                                                   const int* __restrict__ slot,
                                                   double* __restrict__ part) {
  const int t = blockIdx.x;
  double* Tg = tiles + (size_t)target[t] * TB * TB;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int row = 16 * w + li;
  v4d acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int q = src_ptr[t]; q < src_ptr[t + 1]; ++q) {
    const double* Li = tiles + (size_t)src[2 * q] * TB * TB;
    const double* Lk = tiles + (size_t)src[2 * q + 1] * TB * TB;
    double a[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) a[s] = Li[row + TB * (4 * s + lk)];
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(Lk[(16 * c + li) + TB * (4 * s + lk)], a[s],
                                                      acc[c], 0, 0, 0);
  }
  if (slot[t] >= 0) {
    double* Pt = part + (size_t)slot[t] * TB * TB;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) Pt[row + TB * (16 * c + 4 * v + lk)] = acc[c][v];
    return;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) Tg[row + TB * (16 * c + 4 * v + lk)] -= acc[c][v];
}

// the targets split into chunks: target -= the chunks' partial sums, in order
__global__ __launch_bounds__(256) void k_ts_reduce(double* __restrict__ tiles,
                                                   const int* __restrict__ target,
                                                   const int* __restrict__ slot0,
                                                   const int* __restrict__ nslot,
                                                   const double* __restrict__ part) {
  const int r = blockIdx.x;
  double* Tg = tiles + (size_t)target[r] * TB * TB;
  const double* P0 = part + (size_t)slot0[r] * TB * TB;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    double acc = 0.0;
    for (int c = 0; c < nslot[r]; ++c) acc += P0[(size_t)c * TB * TB + e];
    Tg[e] -= acc;
  }
}

// dy := P^T (L L^T)^-1 P g by the tile factor, level by level of the tile
// elimination tree.  Forward: the columns J of a level (one work-group each)
// take x_J := W_J x_J (x_J has every update from the lower levels) and form
// u_p = L_p x_J for their off-diagonal tiles p; then each tile row I the
// level touches subtracts its u_p in column order.  Backward, levels in
// reverse: each column J of a level takes x_J := W_J^T (x_J - sum_p L_p^T
// x_{I(p)}) (the x_I of higher levels are final).  The 64 x 64 products: thread
// t takes row t % 64 over a quarter of the columns, the quarters summed
// through LDS in a fixed order.
__device__ __forceinline__ double ts_matvec(const double* Mt, int trans, const double* xv,
                                            double (*part)[TB]) {
  const int t = threadIdx.x, i = t & 63, q = t >> 6;
  double s = 0.0;
#pragma unroll 4
  for (int k = 16 * q; k < 16 * q + 16; ++k) s += (trans ? Mt[k + TB * i] : Mt[i + TB * k]) * xv[k];
  part[q][i] = s;
  __syncthreads();
  const double v = part[0][i] + part[1][i] + part[2][i] + part[3][i];
  __syncthreads();
  return v;
}

__global__ __launch_bounds__(256) void k_ts_perm(const double* __restrict__ src,
                                                 const int* __restrict__ perm, int m, int n,
                                                 double* __restrict__ dst, int inverse) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  if (!inverse) dst[r] = r < m ? src[perm[r]] : 0.0;
  else if (r < m) dst[perm[r]] = src[r];
}

__global__ __launch_bounds__(256) void k_ts_fwd_cols(const double* __restrict__ tiles,
                                                     const double* __restrict__ winv,
                                                     const int* __restrict__ colptr,
                                                     const int* __restrict__ cols,
                                                     double* __restrict__ x,
                                                     double* __restrict__ u) {
  __shared__ double part[4][TB];
  __shared__ double xv[TB];
  const int J = cols[blockIdx.x], t = threadIdx.x;
  if (t < TB) xv[t] = x[64 * J + t];
  __syncthreads();
  const double y = ts_matvec(winv + (size_t)J * TB * TB, 0, xv, part);
  if (t < TB) {
    x[64 * J + t] = y;
    xv[t] = y;
  }
  __syncthreads();
  for (int p = colptr[J] + 1; p < colptr[J + 1]; ++p) {
    const double v = ts_matvec(tiles + (size_t)p * TB * TB, 0, xv, part);
    if (t < TB) u[(size_t)p * TB + t] = v;
  }
}

__global__ __launch_bounds__(64) void k_ts_fwd_rows(const int* __restrict__ rows,
                                                    const int* __restrict__ src_ptr,
                                                    const int* __restrict__ src,
                                                    const double* __restrict__ u,
                                                    double* __restrict__ x) {
  const int r = blockIdx.x, t = threadIdx.x;
  double acc = 0.0;
  for (int q = src_ptr[r]; q < src_ptr[r + 1]; ++q) acc += u[(size_t)src[q] * TB + t];
  x[64 * rows[r] + t] -= acc;
}

__global__ __launch_bounds__(256) void k_ts_bwd_cols(const double* __restrict__ tiles,
                                                     const double* __restrict__ winv,
                                                     const int* __restrict__ colptr,
                                                     const int* __restrict__ tI,
                                                     const int* __restrict__ cols,
                                                     double* __restrict__ x) {
  __shared__ double part[4][TB];
  __shared__ double xv[TB];
  const int J = cols[blockIdx.x], t = threadIdx.x;
  double acc = 0.0;
  for (int p = colptr[J] + 1; p < colptr[J + 1]; ++p) {
    if (t < TB) xv[t] = x[64 * tI[p] + t];
    __syncthreads();
    acc += ts_matvec(tiles + (size_t)p * TB * TB, 1, xv, part);
  }
  if (t < TB) xv[t] = x[64 * J + t] - acc;
  __syncthreads();
  const double y = ts_matvec(winv + (size_t)J * TB * TB, 1, xv, part);
  if (t < TB) x[64 * J + t] = y;
}

int sparse_schur_upload(TileSparse* ts) {
  auto up = [](int** d, const std::vector<int>& h) -> int {
    SDPA_HIP_CHECK(hipMalloc(d, sizeof(int) * (h.empty() ? 1 : h.size())));
    if (!h.empty())
      SDPA_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return 0;
  };
  if (int e = up(&ts->d_perm, ts->perm)) return e;
  if (int e = up(&ts->d_tI, ts->tI)) return e;
  if (int e = up(&ts->d_tJ, ts->tJ)) return e;
  if (int e = up(&ts->d_colptr, ts->colptr)) return e;
  if (int e = up(&ts->d_lvl_cols, ts->lvl_cols)) return e;
  if (int e = up(&ts->d_trsm_tiles, ts->trsm_tiles)) return e;
  if (int e = up(&ts->d_upd_target, ts->upd_target)) return e;
  if (int e = up(&ts->d_upd_src_ptr, ts->upd_src_ptr)) return e;
  if (int e = up(&ts->d_upd_src, ts->upd_src)) return e;
  if (int e = up(&ts->d_fwd_row, ts->fwd_row)) return e;
  if (int e = up(&ts->d_upd_slot, ts->upd_slot)) return e;
  if (int e = up(&ts->d_red_target, ts->red_target)) return e;
  if (int e = up(&ts->d_red_slot0, ts->red_slot0)) return e;
  if (int e = up(&ts->d_red_nslot, ts->red_nslot)) return e;
  SDPA_HIP_CHECK(hipMalloc(&ts->d_part, sizeof(double) * TB * TB * (size_t)std::max(1, ts->max_slots)));
  if (int e = up(&ts->d_fwd_src_ptr, ts->fwd_src_ptr)) return e;
  if (int e = up(&ts->d_fwd_src, ts->fwd_src)) return e;
  SDPA_HIP_CHECK(hipMalloc(&ts->d_u, sizeof(double) * TB * (size_t)ts->ntiles));
  SDPA_HIP_CHECK(hipMalloc(&ts->d_tiles, sizeof(double) * TB * TB * (size_t)ts->ntiles));
  SDPA_HIP_CHECK(hipMalloc(&ts->d_winv, sizeof(double) * TB * TB * (size_t)ts->T));
  SDPA_HIP_CHECK(hipMalloc(&ts->d_x, sizeof(double) * TB * (size_t)ts->T));
  SDPA_HIP_CHECK(hipMalloc(&ts->d_info, sizeof(int) * (size_t)ts->T));
  return 0;
}

void sparse_schur_free(TileSparse* ts) {
  for (void* p : {(void*)ts->d_perm, (void*)ts->d_tI, (void*)ts->d_tJ, (void*)ts->d_colptr,
                  (void*)ts->d_lvl_cols, (void*)ts->d_trsm_tiles, (void*)ts->d_upd_target,
                  (void*)ts->d_upd_src_ptr, (void*)ts->d_upd_src, (void*)ts->d_tiles,
                  (void*)ts->d_winv, (void*)ts->d_x, (void*)ts->d_info, (void*)ts->d_u,
                  (void*)ts->d_fwd_row, (void*)ts->d_fwd_src_ptr, (void*)ts->d_fwd_src,
                  (void*)ts->d_upd_slot, (void*)ts->d_red_target, (void*)ts->d_red_slot0,
                  (void*)ts->d_red_nslot, (void*)ts->d_part})
    if (p) (void)hipFree(p);
  *ts = TileSparse{};
}

int sparse_schur_factor(hipStream_t s, TileSparse* ts, const double* B, double shift) {
  if (ts->ntiles <= 0) return 0;
  SDPA_HIP_CHECK(hipMemsetAsync(ts->d_info, 0, sizeof(int) * (size_t)ts->T, s));
  hipLaunchKernelGGL(k_ts_gather, dim3(ts->ntiles), dim3(256), 0, s, B, ts->m, ts->d_perm,
                     ts->d_tI, ts->d_tJ, ts->d_tiles, shift);
  for (int l = 0; l < ts->nlev; ++l) {
    const int c0 = ts->lvl_ptr[l], nc = ts->lvl_ptr[l + 1] - c0;
    hipLaunchKernelGGL(k_ts_potrf, dim3(nc), dim3(256), 0, s, ts->d_tiles, ts->d_colptr,
                       ts->d_lvl_cols + c0, ts->d_winv, ts->d_info);
    const int r0 = ts->trsm_ptr[l], nr = ts->trsm_ptr[l + 1] - r0;
    if (nr > 0)
      hipLaunchKernelGGL(k_ts_trsm, dim3(nr), dim3(256), 0, s, ts->d_tiles, ts->d_trsm_tiles + r0,
                         ts->d_tJ, ts->d_winv);
    const int u0 = ts->upd_ptr[l], nu = ts->upd_ptr[l + 1] - u0;
    if (nu > 0)
      hipLaunchKernelGGL(k_ts_update, dim3(nu), dim3(256), 0, s, ts->d_tiles,
                         ts->d_upd_target + u0, ts->d_upd_src_ptr + u0, ts->d_upd_src,
                         ts->d_upd_slot + u0, ts->d_part);
    const int q0 = ts->red_ptr[l], nq = ts->red_ptr[l + 1] - q0;
    if (nq > 0)
      hipLaunchKernelGGL(k_ts_reduce, dim3(nq), dim3(256), 0, s, ts->d_tiles,
                         ts->d_red_target + q0, ts->d_red_slot0 + q0, ts->d_red_nslot + q0,
                         ts->d_part);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  std::vector<int> info(ts->T);
  SDPA_HIP_CHECK(hipMemcpyAsync(info.data(), ts->d_info, sizeof(int) * (size_t)ts->T,
                                hipMemcpyDeviceToHost, s));
  SDPA_HIP_CHECK(hipStreamSynchronize(s));
  for (int J = 0; J < ts->T; ++J)
    if (info[J]) return 64 * J + info[J];
  return 0;
}

int sparse_schur_solve(hipStream_t s, TileSparse* ts, const double* g, double* dy) {
  const int n = 64 * ts->T;
  hipLaunchKernelGGL(k_ts_perm, dim3((n + 255) / 256), dim3(256), 0, s, g, ts->d_perm, ts->m, n,
                     ts->d_x, 0);
  for (int l = 0; l < ts->nlev; ++l) {
    const int c0 = ts->lvl_ptr[l], nc = ts->lvl_ptr[l + 1] - c0;
    hipLaunchKernelGGL(k_ts_fwd_cols, dim3(nc), dim3(256), 0, s, ts->d_tiles, ts->d_winv,
                       ts->d_colptr, ts->d_lvl_cols + c0, ts->d_x, ts->d_u);
    const int r0 = ts->fwd_ptr[l], nr = ts->fwd_ptr[l + 1] - r0;
    if (nr > 0)
      hipLaunchKernelGGL(k_ts_fwd_rows, dim3(nr), dim3(64), 0, s, ts->d_fwd_row + r0,
                         ts->d_fwd_src_ptr + r0, ts->d_fwd_src, ts->d_u, ts->d_x);
  }
  for (int l = ts->nlev - 1; l >= 0; --l) {
    const int c0 = ts->lvl_ptr[l], nc = ts->lvl_ptr[l + 1] - c0;
    hipLaunchKernelGGL(k_ts_bwd_cols, dim3(nc), dim3(256), 0, s, ts->d_tiles, ts->d_winv,
                       ts->d_colptr, ts->d_tI, ts->d_lvl_cols + c0, ts->d_x);
  }
  hipLaunchKernelGGL(k_ts_perm, dim3((n + 255) / 256), dim3(256), 0, s, ts->d_x, ts->d_perm,
                     ts->m, n, dy, 1);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdpa_amd
