// ops.hip -- the device side of the Newton-direction assembly and the
// iterate bookkeeping around the Schur solve, on packed block spaces:
//   A_i . D for every i          InputData::multi_InnerProductToA  (sdpa_dataset.cpp:414-424)
//   sum_i y_i A_i                InputData::multi_plusToA          (sdpa_dataset.cpp:427-433)
//   residuals                    Residuals::compute                (sdpa_dataset.cpp:585-608)
//   r_zinv, gVec                 Newton::compute_rMat/Make_gVec    (sdpa_newton.cpp:793-883)
//   Dz, Dx                       compute_DzMat/compute_DxMat       (sdpa_newton.cpp:2718-2752)
//   step update                  Solutions::update                 (sdpa_dataset.cpp:145-171)
//   X/Z inverse Cholesky         Solutions::computeInverse         (sdpa_dataset.cpp:117-142)
//   inner products for the control scalars (sdpa_parts.cpp).
// Reductions are two-pass with a fixed order, so results are reproducible.
#include "ctx.h"

#include <algorithm>
#include <cmath>

namespace sdpa_amd {

constexpr int kRedChunk = 4096;   // elements per partial sum
constexpr int kMaxChunks = 4096;  // partial slots per reduction

// out_i = sum over the sparse terms of constraint i of v*(D[p1] + D[p2]) (p2 != p1)
__global__ __launch_bounds__(256) void k_adot_sparse(const int* __restrict__ cptr,
                                                     const long* __restrict__ p1,
                                                     const long* __restrict__ p2,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ D, int m,
                                                     double* __restrict__ out) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= m) return;
  double s = 0.0;
  for (int e = cptr[i] + lane; e < cptr[i + 1]; e += 64) {
    double d = D[p1[e]];
    if (p2[e] != p1[e]) d += D[p2[e]];
    s += v[e] * d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[i] = s;
}

// out[con] += dot(pool column, D_q) for every dense member of block q.
__global__ __launch_bounds__(256) void k_adot_dense(const double* __restrict__ pool,
                                                    const int* __restrict__ dlist_con, int n,
                                                    const double* __restrict__ Dq,
                                                    double* __restrict__ out) {
  const size_t nn = (size_t)n * n;
  const double* a = pool + (size_t)blockIdx.x * nn;
  double s = 0.0;
  for (size_t e = threadIdx.x; e < nn; e += 256) s += a[e] * Dq[e];
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out + dlist_con[blockIdx.x], red[0] + red[1] + red[2] + red[3]);
}

// out[pos] = sum_e val_e * y[con_e] over the terms of each distinct packed position.
__global__ void k_yA_sparse(const long* __restrict__ ppos, const long* __restrict__ pptr,
                            const int* __restrict__ pcon, const double* __restrict__ pval,
                            long npos, const double* __restrict__ y, double* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npos) return;
  double s = 0.0;
  for (long e = pptr[t]; e < pptr[t + 1]; ++e) s += pval[e] * y[pcon[e]];
  out[ppos[t]] = s;
}

// Dq[p] += sum_d pool[p, d] * y[con_d]  (dense members of one block)
__global__ void k_yA_dense(const double* __restrict__ pool, const int* __restrict__ dlist_con,
                           int nd, long nn, const double* __restrict__ y, double* __restrict__ Dq) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  double s = 0.0;
  for (int d = 0; d < nd; ++d) s += pool[p + nn * d] * y[dlist_con[d]];
  Dq[p] += s;
}

// generic elementwise: out = a*A + b*B (+ c*Cv)
__global__ void k_axpbypcz(long len, double a, const double* __restrict__ A, double b,
                           const double* __restrict__ B, double cc, const double* __restrict__ Cv,
                           double* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  double r = a * A[t];
  if (B) r += B[t] * b;
  if (Cv) r += Cv[t] * cc;
  out[t] = r;
}

// Two independent k_axpbypcz operations in one launch: operation 0 over
// the first len0 threads, operation 1 over the next len1.
struct Axpb {
  long len;
  double a, b, cc;
  const double* A;
  const double* B;
  const double* Cv;
  double* out;
};

__global__ void k_axpbypcz2(Axpb o0, Axpb o1) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool second = t >= o0.len;
  const Axpb& o = second ? o1 : o0;
  if (second) t -= o0.len;
  if (t >= o.len) return;
  double r = o.a * o.A[t];
  if (o.B) r += o.B[t] * o.b;
  if (o.Cv) r += o.Cv[t] * o.cc;
  o.out[t] = r;
}

// a0 -> b0 and a1 -> b1 (len doubles each) in one launch
__global__ void k_copy2(long len, const double* __restrict__ a0, double* __restrict__ b0,
                        const double* __restrict__ a1, double* __restrict__ b1) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < len) {
    b0[t] = a0[t];
  } else if (t < 2 * len) {
    b1[t - len] = a1[t - len];
  }
}

// The new point in one launch (sdpa_solve.cpp:167-190): X += alphaP Dx,
// Z += alphaD Dz over the packed blocks (len total) and y += alphaD dy
// (len m), each element as k_axpbypcz forms it.
__global__ void k_update_point(long total, long m, double ap, double ad, double* __restrict__ X,
                               const double* __restrict__ Dx, double* __restrict__ Z,
                               const double* __restrict__ Dz, double* __restrict__ y,
                               const double* __restrict__ dy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < total) {
    double rx = 1.0 * X[t];
    rx += Dx[t] * ap;
    double rz = 1.0 * Z[t];
    rz += Dz[t] * ad;
    X[t] = rx;
    Z[t] = rz;
  }
  if (t < m) {
    double ry = 1.0 * y[t];
    ry += dy[t] * ad;
    y[t] = ry;
  }
}

// Lal::getSymmetrize per SDP block: a_ij = a_ji = (a_ij + a_ji) * 0.5
__global__ void k_symmetrize(double* __restrict__ X, const long* __restrict__ off,
                             const int* __restrict__ nn, int nblk) {
  const int q = blockIdx.y;
  if (q >= nblk) return;
  const int n = nn[q];
  double* a = X + off[q];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * n) return;
  const int i = t % n, j = t / n;
  if (i <= j) return;
  const double s = (a[i + (size_t)n * j] + a[j + (size_t)n * i]) * 0.5;
  a[i + (size_t)n * j] = s;
  a[j + (size_t)n * i] = s;
}

// A batch of reductions in two launches: x.y (op 0), max|x| (op 1), min x*y*y
// (op 2, LP step check) of each job, its partial sums over chunks of
// kRedChunk elements (a fixed tree per chunk), then the partials added in
// chunk order: every result is reproducible and independent of the batch.
struct RedJob {
  const double* x;
  const double* y;
  long len;
  int op;
  int out;  // result slot
};
struct RedBatch {
  RedJob j[kMaxRed];
  int n;
};

// grid = (chunks of the longest job, jobs)
__global__ __launch_bounds__(256) void k_reduce_partial(RedBatch B, double* __restrict__ part) {
  const RedJob J = B.j[blockIdx.y];
  const long base = (long)blockIdx.x * kRedChunk;
  if (blockIdx.x > 0 && base >= J.len) return;
  const int op = J.op;
  const double* __restrict__ x = J.x;
  const double* __restrict__ y = J.y;
  double s = (op == 2) ? 1.0e50 : 0.0;
  // the thread's kRedChunk / 256 elements loaded at once, then taken in order
  constexpr int kPer = kRedChunk / 256;
  double xv[kPer], yv[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const long t = base + threadIdx.x + 256 * q;
    xv[q] = t < J.len ? x[t] : 0.0;
    yv[q] = (op != 1 && t < J.len) ? y[t] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    if (base + threadIdx.x + 256 * q >= J.len) break;
    if (op == 0) s += xv[q] * yv[q];
    else if (op == 1) s = fmax(s, fabs(xv[q]));
    else s = fmin(s, xv[q] * yv[q] * yv[q]);
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double o = red[threadIdx.x + w];
      if (op == 0) red[threadIdx.x] += o;
      else if (op == 1) red[threadIdx.x] = fmax(red[threadIdx.x], o);
      else red[threadIdx.x] = fmin(red[threadIdx.x], o);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kMaxChunks + blockIdx.x] = red[0];
}

// thread k: job k's partials in chunk order
__global__ void k_reduce_final(RedBatch B, const double* __restrict__ part,
                               double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= B.n) return;
  const RedJob J = B.j[k];
  const int np = (int)((J.len + kRedChunk - 1) / kRedChunk);
  double s = (J.op == 2) ? 1.0e50 : 0.0;
  for (int i = 0; i < (np > 0 ? np : 1); ++i) {
    const double v = part[(size_t)k * kMaxChunks + i];
    if (J.op == 0) s += v;
    else if (J.op == 1) s = fmax(s, v);
    else s = fmin(s, v);
  }
  out[J.out] = s;
}

// the two launches of a batch (results to out[job.out])
static int reduce_batch(Ctx* c, const RedBatch& B, double* out) {
  if (B.n <= 0) return 0;
  int npmax = 1;
  for (int k = 0; k < B.n; ++k) {
    const long np = (B.j[k].len + kRedChunk - 1) / kRedChunk;
    if (np > kMaxChunks) return -4;
    npmax = np > npmax ? (int)np : npmax;
  }
  hipLaunchKernelGGL(k_reduce_partial, dim3(npmax, B.n), dim3(256), 0, c->stream, B, c->d_red);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, c->stream, B, c->d_red, out);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// sum of the diagonals of the SDP blocks plus the LP entries
// (few blocks: the work-group runs down each diagonal; many: a thread per block)
__global__ void k_trace(const double* __restrict__ X, const long* __restrict__ off,
                        const int* __restrict__ nn, int nblk, long lpOff, int nLP,
                        double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  if (nblk <= 32) {
    for (int q = 0; q < nblk; ++q) {
      const int n = nn[q];
      for (int k = threadIdx.x; k < n; k += 256) s += X[off[q] + (long)k * n + k];
    }
  } else {
    for (int q = threadIdx.x; q < nblk; q += 256) {
      const int n = nn[q];
      const double* Xq = X + off[q];
      for (int k = 0; k < n; ++k) s += Xq[(long)k * n + k];
    }
  }
  for (int t = threadIdx.x; t < nLP; t += 256) s += X[lpOff + t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// LP part of computeInverse: 1/sqrt(x), 1/sqrt(z), 1/z; flag on a negative value
__global__ void k_lp_inverse(const double* __restrict__ X, const double* __restrict__ Z,
                             double* __restrict__ icx, double* __restrict__ icz,
                             double* __restrict__ zi, long lpOff, int nLP, int* __restrict__ bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nLP) return;
  const double x = X[lpOff + t], z = Z[lpOff + t];
  if (x < 0.0 || z < 0.0) atomicOr(bad, 1);
  icx[lpOff + t] = 1.0 / sqrt(x);
  icz[lpOff + t] = 1.0 / sqrt(z);
  zi[lpOff + t] = 1.0 / z;
}

// ---- host wrappers ---------------------------------------------------------

static inline dim3 grid1(long len, int bs = 256) { return dim3((unsigned)((len + bs - 1) / bs)); }


int a_dot(Ctx* c, const double* D, double* out) {
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_adot_sparse, dim3((c->m + 3) / 4), dim3(256), 0, s, c->d_cptr, c->d_cp1,
                     c->d_cp2, c->d_cv, D, c->m, out);
  for (int q = 0; q < c->nSDP; ++q) {
    BlockTable& T = c->blk[q];
    if (T.ndense == 0) continue;
    hipLaunchKernelGGL(k_adot_dense, dim3(T.ndense), dim3(256), 0, s, T.d_dense, T.d_dcon, T.n,
                       D + T.off, out);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

int sum_yA(Ctx* c, const double* y, double* out) {
  hipStream_t s = c->stream;
  SDPA_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(double) * c->total, s));
  if (c->npos > 0)
    hipLaunchKernelGGL(k_yA_sparse, grid1(c->npos), dim3(256), 0, s, c->d_ppos, c->d_pptr,
                       c->d_pcon, c->d_pval, c->npos, y, out);
  for (int q = 0; q < c->nSDP; ++q) {
    BlockTable& T = c->blk[q];
    if (T.ndense == 0) continue;
    long nn = (long)T.n * T.n;
    hipLaunchKernelGGL(k_yA_dense, grid1(nn), dim3(256), 0, s, T.d_dense, T.d_dcon, T.ndense,
                       nn, y, out + T.off);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

static int axpbypcz(Ctx* c, long len, double a, const double* A, double b, const double* B,
                    double cc, const double* Cv, double* out) {
  if (len <= 0) return 0;
  hipLaunchKernelGGL(k_axpbypcz, grid1(len), dim3(256), 0, c->stream, len, a, A, b, B, cc, Cv,
                     out);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

int triple_product(Ctx* c, const double* A, const double* B, const double* Cm, double* R) {
  // R = A*B*C per SDP block (two batched GEMMs through d_w3), LP entries multiply
  const int nb = c->nSDP;
  if (nb > 0) {
    std::vector<GemmItem> it(2 * nb);
    int maxn = 0;
    for (int q = 0; q < nb; ++q) {
      const int n = c->sdpN[q];
      const long o = c->sdpOff[q];
      it[q] = GemmItem{A + o, B + o, c->d_w3 + o, n, n, n, n, n, n};
      it[nb + q] = GemmItem{c->d_w3 + o, Cm + o, R + o, n, n, n, n, n, n};
      maxn = n > maxn ? n : maxn;
    }
    if (nb == 1 && maxn >= kBlasMinN) {
      for (int k = 0; k < 2; ++k) {  // one large block: plain GEMMs through the library
        const GemmItem& g = it[k];
        if (int e = dgemm(c->stream, 0, 0, g.M, g.N, g.K, 1.0, g.A, g.lda, 0, g.B, g.ldb, 0, 0.0, g.C, g.ldc, 0, 1))
          return e;
      }
    } else {
      const void* di = nullptr;
      if (int e = c->dw.desc_copy(c->stream, it.data(), sizeof(GemmItem) * 2 * nb, &di)) return e;
      const GemmItem* d_items = static_cast<const GemmItem*>(di);
      if (int e = gemm_batched(c->stream, d_items, nb, maxn, maxn, 0, 0, 1.0, 0.0, 0))
        return e;
      if (int e = gemm_batched(c->stream, d_items + nb, nb, maxn, maxn, 0, 0, 1.0, 0.0, 0))
        return e;
    }
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_mul3(long len, const double* __restrict__ A, const double* __restrict__ B,
                       const double* __restrict__ C, double* __restrict__ R) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < len) R[t] = A[t] * B[t] * C[t];
}

static int triple(Ctx* c, const double* A, const double* B, const double* Cm, double* R) {
  if (int e = triple_product(c, A, B, Cm, R)) return e;
  if (c->nLP > 0) {
    const long o = c->lpOff;
    hipLaunchKernelGGL(k_mul3, grid1(c->nLP), dim3(256), 0, c->stream, (long)c->nLP, A + o, B + o,
                       Cm + o, R + o);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// p = b - A(X), D = C - sum y A - Z and their max norms into the result slots
// slot, slot + 1, copied to c->h_red[slot..]; no synchronization
int residuals_enqueue(Ctx* c, int slot) {
  if (int e = a_dot(c, c->d_X, c->d_pvec)) return e;
  if (int e = sum_yA(c, c->d_y, c->d_dmat)) return e;
  {
    // p = b - p and D = C - D - Z, one launch
    const Axpb o0{(long)c->m, -1.0, 1.0, 0.0, c->d_pvec, c->d_b, nullptr, c->d_pvec};
    const Axpb o1{(long)c->total, -1.0, 1.0, -1.0, c->d_dmat, c->d_C, c->d_Z, c->d_dmat};
    const long len = o0.len + o1.len;
    if (len > 0)
      hipLaunchKernelGGL(k_axpbypcz2, grid1(len), dim3(256), 0, c->stream, o0, o1);
    SDPA_HIP_CHECK(hipGetLastError());
  }
  RedBatch B;
  B.n = 2;
  B.j[0] = RedJob{c->d_pvec, nullptr, (long)c->m, 1, slot};
  B.j[1] = RedJob{c->d_dmat, nullptr, c->total, 1, slot + 1};
  double* res = c->d_red + (size_t)kMaxRed * kMaxChunks;
  if (int e = reduce_batch(c, B, res)) return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->h_red + slot, res + slot, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
  return 0;
}

int residuals(Ctx* c, double* normP, double* normD) {
  if (int e = residuals_enqueue(c, 0)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  *normP = c->h_red[0];
  *normD = c->h_red[1];
  return 0;
}

// the inner products of codes[0..nq) into the result slots slot0.., copied to
// c->h_red[slot0..]; no synchronization
int dots_enqueue(Ctx* c, int nq, const int* codes, int slot0) {
  hipStream_t s = c->stream;
  if (nq <= 0 || slot0 < 0 || slot0 + nq > kRedResults) return -2;
  double* res = c->d_red + (size_t)kMaxRed * kMaxChunks;
  RedBatch B;
  B.n = 0;
  for (int k = 0; k < nq; ++k) {
    const double *x = nullptr, *y = nullptr;
    long len = c->total;
    switch (codes[k]) {
      case DOT_CX: x = c->d_C; y = c->d_X; break;
      case DOT_BY: x = c->d_b; y = c->d_y; len = c->m; break;
      case DOT_CDX: x = c->d_C; y = c->d_Dx; break;
      case DOT_BDY: x = c->d_b; y = c->d_dy; len = c->m; break;
      case DOT_XDZ: x = c->d_X; y = c->d_Dz; break;
      case DOT_DXZ: x = c->d_Dx; y = c->d_Z; break;
      case DOT_DXDZ: x = c->d_Dx; y = c->d_Dz; break;
      case DOT_XZ: x = c->d_X; y = c->d_Z; break;
      case DOT_X0Z: x = c->d_X0; y = c->d_Z; break;
      case DOT_XZ0: x = c->d_X; y = c->d_Z0; break;
      case DOT_TRX:
      case DOT_TRZ: {
        hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, s, codes[k] == DOT_TRX ? c->d_X : c->d_Z,
                           c->d_sdpOff, c->d_sdpN, c->nSDP, c->lpOff, c->nLP, res + slot0 + k);
        continue;
      }
      default:
        return -2;
    }
    if (!x || !y) return -3;
    B.j[B.n++] = RedJob{x, y, len, 0, slot0 + k};
    if (B.n == kMaxRed) {  // a full batch (the partial slots are reused by the next)
      if (int e = reduce_batch(c, B, res)) return e;
      B.n = 0;
    }
  }
  if (int e = reduce_batch(c, B, res)) return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->h_red + slot0, res + slot0, sizeof(double) * nq,
                                hipMemcpyDeviceToHost, s));
  return 0;
}

int compute_dots(Ctx* c, int nq, const int* codes, double* out) {
  if (int e = dots_enqueue(c, nq, codes, 0)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nq; ++k) out[k] = c->h_red[k];
  return 0;
}

int make_gvec(Ctx* c, int corrector, double target, int infeasible) {
  // r_zinv = target * Zinv - X  (- Dx Dz Zinv in the corrector), and in the
  // infeasible phase w1 = r_zinv - X D Zinv; the products first, then the
  // element-wise steps in one launch (each element formed in the same order
  // of operations as by separate steps)
  if (corrector)
    if (int e = triple(c, c->d_Dx, c->d_Dz, c->d_Zinv, c->d_w1)) return e;
  if (infeasible && !corrector)
    if (int e = triple(c, c->d_X, c->d_dmat, c->d_Zinv, c->d_xrdzinv)) return e;
  const Axpb r{c->total, target, -1.0, -1.0, c->d_Zinv, c->d_X, corrector ? c->d_w1 : nullptr,
               c->d_rzinv};
  Axpb w{0, 0.0, 0.0, 0.0, nullptr, nullptr, nullptr, nullptr};
  const double* D = c->d_rzinv;
  if (infeasible) {
    if (corrector) {
      // needs the finished r_zinv: its own step
      if (int e = axpbypcz(c, c->total, target, c->d_Zinv, -1.0, c->d_X, -1.0, c->d_w1,
                           c->d_rzinv))
        return e;
      if (int e = axpbypcz(c, c->total, 1.0, c->d_rzinv, -1.0, c->d_xrdzinv, 0.0, nullptr,
                           c->d_w1))
        return e;
    } else {
      w = Axpb{c->total, target, -1.0, -1.0, c->d_Zinv, c->d_X, c->d_xrdzinv, c->d_w1};
    }
    D = c->d_w1;
  }
  if (!(infeasible && corrector)) {
    const long len = r.len + w.len;
    if (len > 0) hipLaunchKernelGGL(k_axpbypcz2, grid1(len), dim3(256), 0, c->stream, r, w);
    SDPA_HIP_CHECK(hipGetLastError());
  }
  // g = -A(D) + p
  if (int e = a_dot(c, D, c->d_g)) return e;
  return axpbypcz(c, c->m, -1.0, c->d_g, 1.0, c->d_pvec, 0.0, nullptr, c->d_g);
}

int make_direction(Ctx* c, int infeasible) {
  // Dz = -sum dy A (+ D)
  if (int e = sum_yA(c, c->d_dy, c->d_Dz)) return e;
  if (int e = axpbypcz(c, c->total, -1.0, c->d_Dz, infeasible ? 1.0 : 0.0,
                       infeasible ? c->d_dmat : nullptr, 0.0, nullptr, c->d_Dz))
    return e;
  // Dx = sym(r_zinv - X Dz Zinv)
  if (int e = triple(c, c->d_X, c->d_Dz, c->d_Zinv, c->d_w1)) return e;
  if (int e = axpbypcz(c, c->total, 1.0, c->d_rzinv, -1.0, c->d_w1, 0.0, nullptr, c->d_Dx))
    return e;
  if (c->nSDP > 0) {
    int maxn = 0;
    for (int q = 0; q < c->nSDP; ++q) maxn = c->sdpN[q] > maxn ? c->sdpN[q] : maxn;
    dim3 g((unsigned)(((long)maxn * maxn + 255) / 256), c->nSDP);
    hipLaunchKernelGGL(k_symmetrize, g, dim3(256), 0, c->stream, c->d_Dx, c->d_sdpOff, c->d_sdpN,
                       c->nSDP);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

int update_point(Ctx* c, double alphaP, double alphaD) {
  const long len = std::max<long>(c->total, c->m);
  if (len <= 0) return 0;
  hipLaunchKernelGGL(k_update_point, grid1(len), dim3(256), 0, c->stream, (long)c->total,
                     (long)c->m, alphaP, alphaD, c->d_X, c->d_Dx, c->d_Z, c->d_Dz, c->d_y,
                     c->d_dy);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// The launches of inverse_chol and the copy of its status words (one per
// X/Z block, then the LP flag) to c->h_info; no synchronization.  The
// inverses are formed whether or not a factorization failed (on a failed
// one they hold no meaning; the status says so).
int inverse_chol_enqueue(Ctx* c) {
  // L = chol(X), invCholX = L^-1 ; L = chol(Z), invCholZ = L^-1, Zinv = L^-T L^-1
  const int nb = c->nSDP;
  if (nb > 0) {
    if (c->lpOff > 0)
      hipLaunchKernelGGL(k_copy2, grid1(2 * (long)c->lpOff), dim3(256), 0, c->stream,
                         (long)c->lpOff, c->d_X, c->d_LX, c->d_Z, c->d_LZ);
    std::vector<MatDesc> L(2 * nb), W(2 * nb);
    for (int q = 0; q < nb; ++q) {
      const int n = c->sdpN[q];
      const long o = c->sdpOff[q];
      L[q] = MatDesc{c->d_LX + o, n, n};
      L[nb + q] = MatDesc{c->d_LZ + o, n, n};
      W[q] = MatDesc{c->d_invCholX + o, n, n};
      W[nb + q] = MatDesc{c->d_invCholZ + o, n, n};
    }
    if (int e = potrf_batched(c->stream, &c->dw, L.data(), 2 * nb, 0, c->d_info)) return e;
    // reuse the 64x64 diagonal inverses the factorization left in dw.d_dinv
    if (int e = trtri_batched(c->stream, &c->dw, L.data(), W.data(), 2 * nb, c->dw.d_dinv))
      return e;
    // Zinv = invCholZ^T * invCholZ
    std::vector<GemmItem> it(nb);
    int maxn = 0;
    for (int q = 0; q < nb; ++q) {
      const int n = c->sdpN[q];
      const long o = c->sdpOff[q];
      it[q] = GemmItem{c->d_invCholZ + o, c->d_invCholZ + o, c->d_Zinv + o, n, n, n, n, n, n};
      maxn = n > maxn ? n : maxn;
    }
    if (nb == 1 && maxn >= kBlasMinN) {
      const GemmItem& g = it[0];  // one large block: a plain GEMM through the library
      if (int e = dgemm(c->stream, 1, 0, g.M, g.N, g.K, 1.0, g.A, g.lda, 0, g.B, g.ldb, 0, 0.0, g.C, g.ldc, 0, 1))
        return e;
    } else {
      const void* di = nullptr;
      if (int e = c->dw.desc_copy(c->stream, it.data(), sizeof(GemmItem) * nb, &di)) return e;
      if (int e = gemm_batched(c->stream, static_cast<const GemmItem*>(di), nb, maxn, maxn, 1, 0,
                               1.0, 0.0, 0))
        return e;
    }
  }
  if (c->nLP > 0) {
    int* bad = c->d_info + 2 * nb;
    SDPA_HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_lp_inverse, grid1(c->nLP), dim3(256), 0, c->stream, c->d_X, c->d_Z,
                       c->d_invCholX, c->d_invCholZ, c->d_Zinv, c->lpOff, c->nLP, bad);
  }
  const int nstat = 2 * nb + (c->nLP > 0 ? 1 : 0);
  if (nstat > 0)
    SDPA_HIP_CHECK(hipMemcpyAsync(c->h_info, c->d_info, sizeof(int) * nstat,
                                  hipMemcpyDeviceToHost, c->stream));
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// after the stream has passed inverse_chol_enqueue: 1 if a factorization failed
int inverse_chol_status(const Ctx* c) {
  const int nstat = 2 * c->nSDP + (c->nLP > 0 ? 1 : 0);
  for (int b = 0; b < nstat; ++b)
    if (c->h_info[b] != 0) return 1;
  return 0;
}

int inverse_chol(Ctx* c) {
  if (int e = inverse_chol_enqueue(c)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return inverse_chol_status(c);
}

// min over the LP coordinates of Dx * L_X^2 and Dz * L_Z^2 (Jal::getMinEigen's LP
// part, sdpa_jordan.cpp:101-106) into the result slots slot, slot + 1, copied
// to c->h_red[slot..]; no synchronization (nLP > 0)
int lp_min_enqueue(Ctx* c, int slot) {
  RedBatch B;
  B.n = 2;
  B.j[0] = RedJob{c->d_Dx + c->lpOff, c->d_invCholX + c->lpOff, (long)c->nLP, 2, slot};
  B.j[1] = RedJob{c->d_Dz + c->lpOff, c->d_invCholZ + c->lpOff, (long)c->nLP, 2, slot + 1};
  double* res = c->d_red + (size_t)kMaxRed * kMaxChunks;
  if (int e = reduce_batch(c, B, res)) return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->h_red + slot, res + slot, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
  return 0;
}

}  // namespace sdpa_amd
