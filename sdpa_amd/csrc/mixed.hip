// mixed.hip -- the mixed-precision Schur factorization and the FP64
// iterative refinement of the Schur solves (the FP32 factor is never used
// alone: every solution is refined against the FP64 Schur matrix until the
// residual is at FP64 accuracy, or the refinement is reported as stalled and
// the caller factors in FP64).  No counterpart in the reference, which
// factors bMat in FP64 only (sdpa_linear.cpp:762-790); the solve it
// replaces is Lal::solveSystems (sdpa_linear.cpp:843-862).
#include "ctx.h"

#include <cmath>

namespace sdpa_amd {

constexpr int kMaxRefine = 10;

static int mixed_fail(Ctx* c, int code, const char* msg) {
  c->last_error = msg;
  return code;
}

void free_mixed(Ctx* c) {
  for (void* p : {(void*)c->d_B32, (void*)c->d_scale, (void*)c->d_P64, (void*)c->d_D64,
                  (void*)c->d_mr, (void*)c->d_mx, (void*)c->d_md, (void*)c->d_symv,
                  (void*)c->d_mnorm, (void*)c->d_mflags, (void*)c->d_bmax})
    if (p) (void)hipFree(p);
  if (c->h_mnorm) (void)hipHostFree(c->h_mnorm);
  if (c->h_mflags) (void)hipHostFree(c->h_mflags);
  c->d_B32 = nullptr;
  c->d_scale = c->d_P64 = c->d_D64 = c->d_mr = c->d_mx = c->d_md = c->d_symv = nullptr;
  c->d_mnorm = nullptr;
  c->h_mnorm = nullptr;
  c->d_mflags = c->h_mflags = nullptr;
  c->d_bmax = nullptr;
  c->mixed_m = 0;
  c->factor_mixed = 0;
}

static int ensure_mixed(Ctx* c) {
  const int m = c->m;
  if (c->mixed_m == m) return 0;
  free_mixed(c);
  const size_t nt = (m + 63) / 64, nb = (m + 255) / 256;
  SDPA_HIP_CHECK(hipMalloc(&c->d_B32, sizeof(float) * (size_t)m * m));
  SDPA_HIP_CHECK(hipMalloc(&c->d_scale, sizeof(double) * m));
  SDPA_HIP_CHECK(hipMalloc(&c->d_P64, sizeof(double) * (size_t)m * 256));
  SDPA_HIP_CHECK(hipMalloc(&c->d_D64, sizeof(double) * nb * 65536));
  SDPA_HIP_CHECK(hipMalloc(&c->d_mr, sizeof(double) * m));
  SDPA_HIP_CHECK(hipMalloc(&c->d_mx, sizeof(double) * m));
  SDPA_HIP_CHECK(hipMalloc(&c->d_md, sizeof(double) * m));
  SDPA_HIP_CHECK(hipMalloc(&c->d_symv, sizeof(double) * nt * nt * 64));
  SDPA_HIP_CHECK(hipMalloc(&c->d_mnorm, sizeof(double) * 4));
  SDPA_HIP_CHECK(hipMalloc(&c->d_mflags, sizeof(int) * 2));
  SDPA_HIP_CHECK(hipMalloc(&c->d_bmax, sizeof(unsigned long long)));
  SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_mnorm, sizeof(double) * 4, hipHostMallocDefault));
  SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_mflags, sizeof(int) * 2, hipHostMallocDefault));
  c->mixed_m = m;
  return 0;
}

// partials of y = B x (ABS: |B| x) with B given by its lower triangle: 64x64
// tile (I, J), I >= J, adds B_IJ x_J to y_I (slot (I, J)) and, below the
// diagonal, B_IJ^T x_I to y_J (slot (J, I)); a diagonal tile puts both in
// (I, I).  grid (nt, nt): the tiles above the diagonal return at once.
template <bool ABS>
__global__ __launch_bounds__(256) void k_symv_tiles(const double* __restrict__ B, int ld, int m,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ part, int nt) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (J > I) return;
  __shared__ double red[4][64];
  __shared__ double tr[64][65];
  const int t = threadIdx.x, li = t & 63, cg = t >> 6;
  const int r = I * 64 + li, rr = min(r, m - 1);
  const double xr = r < m ? x[r] : 0.0;
  double v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = B[rr + (size_t)ld * min(J * 64 + 16 * cg + k, m - 1)];
    if (ABS) v[k] = fabs(v[k]);
  }
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = J * 64 + 16 * cg + k;
    const bool ok = r < m && c < m;
    const double xc = c < m ? x[c] : 0.0;
    if (ok && (I > J || c <= r)) d = fma(v[k], xc, d);
    tr[16 * cg + k][li] = (ok && (I > J || c < r)) ? v[k] * xr : 0.0;
  }
  red[cg][li] = d;
  __syncthreads();
  if (t < 64) {
    double y = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    if (I == J) {
      double u = 0.0;
      for (int l = 0; l < 64; ++l) u += tr[t][l];
      y += u;
    }
    part[((size_t)I * nt + J) * 64 + t] = y;
  } else if (t < 128 && I > J) {
    const int c = t - 64;
    double u = 0.0;
    for (int l = 0; l < 64; ++l) u += tr[c][l];
    part[((size_t)J * nt + I) * 64 + c] = u;
  }
}

// r = g - y (y itself when g is null), y_i the sum of the partials of row
// block i / 64 in a fixed order: thread (row, q) sums the partials
// S = q (mod 4), 4 loads in flight, then the four quarters are added
__global__ __launch_bounds__(256) void k_symv_finish(const double* __restrict__ part, int nt,
                                                     int m, const double* __restrict__ g,
                                                     double* __restrict__ r) {
  __shared__ double red[4][64];
  const int R = blockIdx.x, t = threadIdx.x, li = t & 63, q = t >> 6;
  const double* pp = part + (size_t)R * nt * 64 + li;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int S = q;
  for (; S + 12 < nt; S += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] += pp[(size_t)(S + 4 * u) * 64];
  }
  for (; S < nt; S += 4) a[0] += pp[(size_t)S * 64];
  red[q][li] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  const int i = R * 64 + t;
  if (t < 64 && i < m) {
    const double y = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    r[i] = g ? g[i] - y : y;
  }
}

__global__ __launch_bounds__(256) void k_mx_scale(double* __restrict__ dst,
                                                  const double* __restrict__ src,
                                                  const double* __restrict__ scale, int m) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < m) dst[i] = src[i] * scale[i];
}

// x = (first ? 0 : x) + S d
__global__ __launch_bounds__(256) void k_mx_accum(double* __restrict__ x,
                                                  const double* __restrict__ d,
                                                  const double* __restrict__ scale, int m,
                                                  int first) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < m) x[i] = (first ? 0.0 : x[i]) + scale[i] * d[i];
}

// norms of the equilibrated system (S B S) (S^-1 x) = S g, whose matrix has
// a unit diagonal: out = (max |s_i r_i|, max |x_i / s_i|, max |s_i g_i|),
// one work-group
__global__ __launch_bounds__(1024) void k_mx_norms(const double* __restrict__ r,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ g,
                                                   const double* __restrict__ scale, int m,
                                                   double* __restrict__ out) {
  __shared__ double sh[3][16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = t; i < m; i += 1024) {
    const double si = scale[i];
    a = fmax(a, fabs(si * r[i]));
    b = fmax(b, fabs(x[i] / si));
    c = fmax(c, fabs(si * g[i]));
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o, 64));
    b = fmax(b, __shfl_xor(b, o, 64));
    c = fmax(c, __shfl_xor(c, o, 64));
  }
  if (lane == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
  __syncthreads();
  if (t == 0) {
    double A = 0.0, Bs = 0.0, C = 0.0;
    for (int k = 0; k < 16; ++k) { A = fmax(A, sh[0][k]); Bs = fmax(Bs, sh[1][k]); C = fmax(C, sh[2][k]); }
    out[0] = A;
    out[1] = Bs;
    out[2] = C;
  }
}

// ||S B S||_inf = max_i s_i (|B| s)_i, y = |B| s given; one work-group
__global__ __launch_bounds__(1024) void k_mx_anorm(const double* __restrict__ y,
                                                   const double* __restrict__ scale, int m,
                                                   double* __restrict__ out) {
  __shared__ double sh[16];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int i = t; i < m; i += 1024) a = fmax(a, scale[i] * y[i]);
  for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
  if ((t & 63) == 0) sh[t >> 6] = a;
  __syncthreads();
  if (t == 0) {
    double A = 0.0;
    for (int k = 0; k < 16; ++k) A = fmax(A, sh[k]);
    *out = A;
  }
}

int factor_mixed(Ctx* c) {
  const int m = c->m;
  // outside the float solves' range (m <= 512, or more than 48 blocks of 256
  // rows): the FP64 factorization
  if (m <= 512 || !chol_solve_f32_supported(m, m)) return 1;
  if (int e = ensure_mixed(c)) return e;
  hipStream_t s = c->stream;
  if (int e = potrf_mixed(s, &c->dw, c->d_B, m, m, c->d_B32, c->d_scale, c->d_P64, c->d_D64,
                          c->d_Binv, c->d_mflags, c->d_mflags + 1, c->d_bmax))
    return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->h_mflags, c->d_mflags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  SDPA_HIP_CHECK(hipStreamSynchronize(s));
  if (c->h_mflags[0] != 0 || c->h_mflags[1] != 0) return 1;
  // ||S B S||_inf for the refinement's stopping test
  {
    const int nt = (m + 63) / 64;
    hipLaunchKernelGGL(k_symv_tiles<true>, dim3(nt, nt), dim3(256), 0, s, c->d_B, m, m,
                       c->d_scale, c->d_symv, nt);
    hipLaunchKernelGGL(k_symv_finish, dim3(nt), dim3(256), 0, s, c->d_symv, nt, m,
                       (const double*)nullptr, c->d_md);
    hipLaunchKernelGGL(k_mx_anorm, dim3(1), dim3(1024), 0, s, c->d_md, c->d_scale, m,
                       c->d_mnorm + 3);
    SDPA_HIP_CHECK(hipMemcpyAsync(c->h_mnorm + 3, c->d_mnorm + 3, sizeof(double),
                                  hipMemcpyDeviceToHost, s));
  }
  // inverses of the FP64 256x256 diagonal blocks, from the 64x64 inverses of
  // the panel factorizations
  const int nb = (m + 255) / 256;
  std::vector<MatDesc> Ls(nb), Ws(nb);
  for (int b = 0; b < nb; ++b) {
    const int kb = std::min(256, m - 256 * b);
    Ls[b] = MatDesc{c->d_D64 + (size_t)b * 65536, kb, 256};
    Ws[b] = MatDesc{c->d_Winv + (size_t)b * 65536, kb, 256};
  }
  if (int e = trtri_batched(s, &c->dw, Ls.data(), Ws.data(), nb, c->d_Binv)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(s));
  c->mx_anorm = c->h_mnorm[3];
  c->factor_mixed = 1;
  ++c->mx_factors;
  return 0;
}

int solve_mixed(Ctx* c) {
  const int m = c->m, nt = (m + 63) / 64;
  hipStream_t s = c->stream;
  const dim3 vg((m + 255) / 256);
  SDPA_HIP_CHECK(hipMemcpyAsync(c->d_mr, c->d_g, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
  ++c->mx_solves;
  double prev = HUGE_VAL;
  for (int it = 0; it < kMaxRefine; ++it) {
    // correction d = S (L L^T)^-1 S r, x += d
    hipLaunchKernelGGL(k_mx_scale, vg, dim3(256), 0, s, c->d_md, c->d_mr, c->d_scale, m);
    if (int e = chol_solve_f32(s, &c->dw, c->d_B32, m, m, c->d_Winv, c->d_md))
      return e > 0 ? mixed_fail(c, -33, "mixed solve: pipelined solves unavailable") : e;
    hipLaunchKernelGGL(k_mx_accum, vg, dim3(256), 0, s, c->d_mx, c->d_md, c->d_scale, m,
                       it == 0 ? 1 : 0);
    // residual r = g - B x against the FP64 Schur matrix
    hipLaunchKernelGGL(k_symv_tiles<false>, dim3(nt, nt), dim3(256), 0, s, c->d_B, m, m, c->d_mx,
                       c->d_symv, nt);
    hipLaunchKernelGGL(k_symv_finish, dim3(nt), dim3(256), 0, s, c->d_symv, nt, m, c->d_g,
                       c->d_mr);
    hipLaunchKernelGGL(k_mx_norms, dim3(1), dim3(1024), 0, s, c->d_mr, c->d_mx, c->d_g,
                       c->d_scale, m, c->d_mnorm);
    SDPA_HIP_CHECK(hipMemcpyAsync(c->h_mnorm, c->d_mnorm, sizeof(double) * 3,
                                  hipMemcpyDeviceToHost, s));
    SDPA_HIP_CHECK(hipStreamSynchronize(s));
    if (c->dw.h_err && *c->dw.h_err) {
      if (int e = c->dw.clear_err(s)) return e;
      return mixed_fail(c, -60, "pipelined solve gave up waiting");
    }
    ++c->mx_refine;
    const double nr = c->h_mnorm[0], nx = c->h_mnorm[1];
    // FP64 accuracy on the equilibrated system, the test of LAPACK's mixed
    // precision dsposv: ||r|| < sqrt(m) eps ||A||_inf ||x|| (all infinity
    // norms, eps = 2^-53)
    const double tol = std::sqrt((double)m) * 1.1102230246251565e-16 * c->mx_anorm * nx;
    if (!(nr == nr)) break;  // NaN: stalled
    if (nr <= tol) {
      SDPA_HIP_CHECK(hipMemcpyAsync(c->d_dy, c->d_mx, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
      return 0;
    }
    if (it > 0 && nr > 0.5 * prev) break;
    prev = nr;
  }
  ++c->mx_stalls;
  return 2;
}

}  // namespace sdpa_amd
