// gemm.hip -- FP64 and FP32 matrix products on CDNA4 MFMA (gfx950),
// strided-batched: C_b = alpha op(A_b) op(B_b) + beta C_b, column-major,
// op = N or T (FP64: v_mfma_f64_16x16x4f64; FP32: v_mfma_f32_16x16x4f32 with
// FP32 accumulation, the same tiling).  Used for
// the dense products of the hot path: the F1/F2 products X A_j and
// (X A_j) Z^-1 of dense constraint blocks, the long-inner-dimension product
// Abar^T Gbar (split over slices of the inner dimension by the caller, one
// batch item per slice), Z^-1 = L^-T L^-1 and L D L^T for single large
// blocks.
//
// One work-group of 256 threads (4 waves) computes a 128 x 128 tile of C; each
// wave a 64 x 64 quarter as 4 x 4 tiles of v_mfma_f64_16x16x4f64 (16
// accumulators); small products take 64 x 64 tiles (see Tile).  The operands
// go through LDS in k-slices (two buffers): each thread loads its 8 + 8
// elements of the next slice into registers while the waves multiply the
// current one, then writes them to the other buffer -- one barrier per
// slice.  Each operand keeps in LDS the
// order in which it is contiguous in memory: k-major (As[k][m]) when m runs
// along memory (op = N for A), m-major with a 17-double row (As[m][k]) when
// k does; both the stores and a wave's MFMA fragment reads (16 m at 4 k)
// are then free of bank conflicts.  Edges: loads beyond M, N or K read zero,
// stores are masked.
#include "ctx.h"

namespace sdpa_amd {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T>
struct GemmArgs {
  const T* A;
  const T* B;
  T* C;
  int M, N, K, lda, ldb, ldc;
  long sA, sB, sC;
  T alpha, beta;
};

// one 16 x 16 x 4 MFMA step and its accumulator type per element type
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc;
  static __device__ __forceinline__ acc step(double x, double y, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef v4f acc;
  static __device__ __forceinline__ acc step(float x, float y, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
};

// Tile shapes: BT x BT output tile per work-group, k-slices of BK.  Large
// products take 128 x 128 tiles with 16-deep slices (each wave a 64 x 64
// quarter: 16 MFMA accumulators); products with few such tiles take 64 x 64
// tiles with 32-deep slices (a 32 x 32 quarter per wave, half the slices per
// k), so that more work-groups run and each waits on fewer slice loads.
template <int BT, int BK>
struct Tile {
  static constexpr int LDK = BT + 2;         // k-major slices: a k row of BT (pad)
  static constexpr int LDM = BK + 1;         // m-major slices: an m row of BK (pad)
  static constexpr int SL = BT * LDM > BK * LDK ? BT * LDM : BK * LDK;  // doubles per slice
  static constexpr int PER = BT * BK / 256;  // elements per thread per operand
  static constexpr int NA = BT / 32;         // MFMA tiles per wave side
};

// Slice coordinates of thread t's q-th element of op(A) (or op(B)): for
// op = N the BT rows of one k are contiguous in memory (consecutive threads
// take consecutive m), for op = T the BK k of one row are.
template <int T, int BT, int BK>
__device__ __forceinline__ void slice_pos(int e, int& mm, int& kk) {
  if (T == 0) { mm = e % BT; kk = e / BT; }
  else { kk = e % BK; mm = e / BK; }
}

// LDS position of op(A)'s element (m, k) of a slice (op(B)'s (n, k)): each
// operand keeps the order in which it is contiguous in memory.
template <int T, int BT, int BK>
__device__ __forceinline__ int lds_pos(int mm, int kk) {
  return T == 0 ? kk * Tile<BT, BK>::LDK + mm : mm * Tile<BT, BK>::LDM + kk;
}

template <class T, int TA, int TB, int BT, int BK>
__device__ __forceinline__ void load_slice(const GemmArgs<T>& g, const T* A, const T* B,
                                           int m0, int n0, int k0, T (&ra)[Tile<BT, BK>::PER],
                                           T (&rb)[Tile<BT, BK>::PER]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Tile<BT, BK>::PER; ++q) {
    const int e = t + 256 * q;
    int mm, kk;
    slice_pos<TA, BT, BK>(e, mm, kk);
    const int gm = m0 + mm, gk = k0 + kk;
    const bool in = gm < g.M && gk < g.K;
    ra[q] = in ? (TA == 0 ? A[gm + (size_t)g.lda * gk] : A[gk + (size_t)g.lda * gm]) : T(0);
    // op(B)'s "m" is n; op(B) = B (K x N, k contiguous) is the transposed
    // case of the pattern, op(B) = B^T (B is N x K, n contiguous) the direct one
    int nn, kb;
    slice_pos<1 - TB, BT, BK>(e, nn, kb);
    const int gn = n0 + nn, gkb = k0 + kb;
    const bool inb = gn < g.N && gkb < g.K;
    rb[q] = inb ? (TB == 0 ? B[gkb + (size_t)g.ldb * gn] : B[gn + (size_t)g.ldb * gkb]) : T(0);
  }
}

template <class T, int TA, int TB, int BT, int BK>
__device__ __forceinline__ void store_slice(const T (&ra)[Tile<BT, BK>::PER],
                                            const T (&rb)[Tile<BT, BK>::PER], T* As, T* Bs) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Tile<BT, BK>::PER; ++q) {
    const int e = t + 256 * q;
    int mm, kk;
    slice_pos<TA, BT, BK>(e, mm, kk);
    As[lds_pos<TA, BT, BK>(mm, kk)] = ra[q];
    int nn, kb;
    slice_pos<1 - TB, BT, BK>(e, nn, kb);
    Bs[lds_pos<1 - TB, BT, BK>(nn, kb)] = rb[q];
  }
}

template <class T, int TA, int TB, int BT, int BK>
__global__ __launch_bounds__(256, 2) void k_gemm_mfma(GemmArgs<T> g) {
  using TL = Tile<BT, BK>;
  using Acc = typename Mfma<T>::acc;
  constexpr int NA = TL::NA, WT = BT / 2;
  __shared__ T As[2][TL::SL];
  __shared__ T Bs[2][TL::SL];
  const int bt = blockIdx.z;
  const T* A = g.A + g.sA * bt;
  const T* B = g.B + g.sB * bt;
  T* C = g.C + g.sC * bt;
  const int tilesM = (g.M + BT - 1) / BT;
  const int m0 = (blockIdx.x % tilesM) * BT, n0 = (blockIdx.x / tilesM) * BT;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int wm = WT * (w & 1), wn = WT * (w >> 1);  // the wave's quarter
  Acc acc[NA][NA];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NA; ++b) acc[a][b] = Acc{T(0), T(0), T(0), T(0)};
  const int nk = (g.K + BK - 1) / BK;
  T ra[TL::PER], rb[TL::PER];
  load_slice<T, TA, TB, BT, BK>(g, A, B, m0, n0, 0, ra, rb);
  store_slice<T, TA, TB, BT, BK>(ra, rb, As[0], Bs[0]);
  __syncthreads();
  for (int ks = 0; ks < nk; ++ks) {
    const int cur = ks & 1;
    if (ks + 1 < nk) load_slice<T, TA, TB, BT, BK>(g, A, B, m0, n0, (ks + 1) * BK, ra, rb);
    const T* as = As[cur];
    const T* bs = Bs[cur];
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      T y[NA], x[NA];
#pragma unroll
      for (int a = 0; a < NA; ++a) y[a] = as[lds_pos<TA, BT, BK>(wm + 16 * a + li, k4 + lk)];
#pragma unroll
      for (int b = 0; b < NA; ++b) x[b] = bs[lds_pos<1 - TB, BT, BK>(wn + 16 * b + li, k4 + lk)];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b)
          acc[a][b] = Mfma<T>::step(x[b], y[a], acc[a][b]);
    }
    if (ks + 1 < nk) {
      store_slice<T, TA, TB, BT, BK>(ra, rb, As[cur ^ 1], Bs[cur ^ 1]);
      __syncthreads();
    }
  }
  // FP64: C[m0 + wm + 16 a + li][n0 + wn + 16 b + 4 v + lk]; FP32: the
  // accumulator's v-th value of lane (li, lk) is column 4 lk + v
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int gm = m0 + wm + 16 * a + li;
    if (gm >= g.M) continue;
#pragma unroll
    for (int b = 0; b < NA; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gn = n0 + wn + 16 * b + (sizeof(T) == 8 ? 4 * v + lk : 4 * lk + v);
        if (gn < g.N) {
          T* p = C + gm + (size_t)g.ldc * gn;
          const T r = g.alpha * acc[a][b][v];
          *p = (g.beta == T(0)) ? r : r + g.beta * *p;
        }
      }
  }
}

template <class T, int BT, int BK>
static void launch(hipStream_t s, int ta, int tb, const GemmArgs<T>& g, dim3 grid) {
  if (ta == 0 && tb == 0) hipLaunchKernelGGL((k_gemm_mfma<T, 0, 0, BT, BK>), grid, dim3(256), 0, s, g);
  else if (ta == 0) hipLaunchKernelGGL((k_gemm_mfma<T, 0, 1, BT, BK>), grid, dim3(256), 0, s, g);
  else if (tb == 0) hipLaunchKernelGGL((k_gemm_mfma<T, 1, 0, BT, BK>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_gemm_mfma<T, 1, 1, BT, BK>), grid, dim3(256), 0, s, g);
}

template <class T>
static int gemm_any(hipStream_t s, int ta, int tb, int M, int N, int K, T alpha, const T* A,
                    int lda, long sA, const T* B, int ldb, long sB, T beta, T* C, int ldc, long sC,
                    int batch) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  if (batch > 65535) return -43;
  GemmArgs<T> g{A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC, alpha, beta};
  // 128-wide tiles when they give the chip at least two work-groups per CU
  // (512), else 64-wide ones
  const long t128 = (long)((M + 127) / 128) * ((N + 127) / 128);
  // FP32 slices are twice as deep (the same LDS bytes): its MFMA runs at
  // about three times the FP64 rate, so each slice load must feed more of it
  constexpr int kDeep = sizeof(T) == 8 ? 1 : 2;
  if (t128 * batch >= 512) {
    launch<T, 128, 16 * kDeep>(s, ta, tb, g, dim3((unsigned)t128, 1, (unsigned)batch));
  } else {
    const long t64 = (long)((M + 63) / 64) * ((N + 63) / 64);
    launch<T, 64, 32 * kDeep>(s, ta, tb, g, dim3((unsigned)t64, 1, (unsigned)batch));
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

int dgemm(hipStream_t s, int ta, int tb, int M, int N, int K, double alpha, const double* A,
          int lda, long sA, const double* B, int ldb, long sB, double beta, double* C, int ldc,
          long sC, int batch) {
  return gemm_any<double>(s, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                          batch);
}

int sgemm(hipStream_t s, int ta, int tb, int M, int N, int K, float alpha, const float* A,
          int lda, long sA, const float* B, int ldb, long sB, float beta, float* C, int ldc,
          long sC, int batch) {
  return gemm_any<float>(s, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                         batch);
}

}  // namespace sdpa_amd
