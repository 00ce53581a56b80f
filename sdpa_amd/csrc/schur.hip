// schur.hip -- Schur complement matrix bMat on the device
// (Newton::Make_bMat dense path, sdpa_newton.cpp:2634-2670).
//
//   bMat[i,j] = sum over SDP blocks of A_i . (X A_j Z^-1)  +  LP term,
//
// written into the lower triangle (i >= j) of the m x m column-major buffer;
// the Cholesky and the solves read only that triangle.  Every unordered pair
// is computed once, by its owner j (more NonZeroEffect; ties go to the smaller
// index, sdpa_newton.cpp:2079-2081) with the owner's formula:
//   F3 (sparse owner, calF3_thread_2, 1319-1403): gather over the two nonzero
//      lists.  Light pairs (both lists <= kLight entries) are one thread each;
//      pairs that involve a heavy list are one wave each with a wave reduction.
//   F1/F2 (calF1/calF2, 885-936): G_j = X A_j Z^-1 by batched MFMA GEMMs, then
//      A_i . G_j -- one GEMM (K = n^2) for dense partners, a gather for
//      sparse ones.  F2 is evaluated through G_j as well (same value).
//   LP (compute_bMat_dense_LP, 2562-2602): x_l z_l^-1 a_il a_jl.
#include "ctx.h"

namespace sdpa_amd {

constexpr int kLight = 8;

// Entry (u, v) of the symmetric n x n matrix M (column-major): read as
// M[v + n u] (ByOwner = false: rows u of the owner's entries are the
// columns) or, through the symmetry, as M[u + n v] (ByOwner = true: when the
// lanes of a wave take consecutive owner entries, their reads of one partner
// column are contiguous).
template <bool ByOwner>
__device__ __forceinline__ double sym_at(const double* __restrict__ M, int n, int u, int v) {
  return ByOwner ? M[u + (size_t)n * v] : M[v + (size_t)n * u];
}

template <bool ByOwner = false>
__device__ __forceinline__ double f3_pair(const int* __restrict__ er, const int* __restrict__ ec,
                                          const double* __restrict__ ev, int so, int co, int sp,
                                          int cp, int n, const double* __restrict__ X,
                                          const double* __restrict__ Zi, int e1_begin,
                                          int e1_step) {
  double ret = 0.0;
  for (int e1 = e1_begin; e1 < co; e1 += e1_step) {
    const int a = er[so + e1], b = ec[so + e1];
    const double v1 = ev[so + e1];
    double sum = 0.0;
    for (int e2 = 0; e2 < cp; ++e2) {
      const int g = er[sp + e2], d = ec[sp + e2];
      // calF3's four terms: Zb[d] Xa[g] + Zb[g] Xa[d] + Za[d] Xb[g] + Za[g] Xb[d]
      const double t1 = sym_at<ByOwner>(Zi, n, b, d) * sym_at<ByOwner>(X, n, a, g);
      double plu;
      if (a != b) {
        if (g != d)
          plu = t1 + sym_at<ByOwner>(Zi, n, b, g) * sym_at<ByOwner>(X, n, a, d) +
                sym_at<ByOwner>(Zi, n, a, d) * sym_at<ByOwner>(X, n, b, g) +
                sym_at<ByOwner>(Zi, n, a, g) * sym_at<ByOwner>(X, n, b, d);
        else
          plu = t1 + sym_at<ByOwner>(Zi, n, a, d) * sym_at<ByOwner>(X, n, b, g);
      } else {
        if (g != d) plu = t1 + sym_at<ByOwner>(Zi, n, b, g) * sym_at<ByOwner>(X, n, a, d);
        else plu = t1;
      }
      sum += ev[sp + e2] * plu;
    }
    ret += v1 * sum;
  }
  return ret;
}

// k_f3_multi: the light F3 pairs of every SDP block of n <= kMultiN in one
// launch, one work-group per Schur column j, the column kept in LDS (acc).
//
// With c = v (a != b) or v / 2 (a == b) for an entry (a, b) of value v, the
// value calF3 (sdpa_newton.cpp:1319-1403) gives the entry pair (a, b), (g, d)
// is c c' (X[a][g] Z[b][d] + X[a][d] Z[b][g] + X[b][g] Z[a][d] + X[b][d] Z[a][g])
// with c' the partner's (f3_pair's four terms, its two-term forms being the
// four-term one halved for a == b and for g == d; Z here is Z^-1).  It is
// symmetric in (a, b) and in (g, d).  A visit of column j to a block
// therefore first forms, for every position p = (lo, hi), lo <= hi, of the
// block's upper triangle,
//   R[p] = s_p sum_e (U1_e[lo] W1_e[hi] + U1_e[hi] W1_e[lo] + U2_e[lo] W2_e[hi] + U2_e[hi] W2_e[lo])
// over A_j's entries e = (a, b) in the block, U1 = c X[a][.], W1 = Z[b][.],
// U2 = c X[b][.], W2 = Z[a][.] (rows staged in LDS), s_p = 1 (lo != hi) or
// 1/2; then each partner i >= j of the block's list adds sum over its
// entries of v' R[p'] to acc[i]: one LDS read per entry pair.  The pair is
// evaluated here when its owner (more NonZeroEffect, ties to the smaller
// index, sdpa_newton.cpp:2079-2081) takes F3 and both lists are light; heavy
// lists are neither visited nor listed (k_f3_heavy*).
//
// Entry records (16 bytes): meta = i | p << 13 | (entries of i after this
// one) << 25 | head << 28 | (i takes F3) << 29, i's NonZeroEffect, v'.  A
// block's records come in chunks of 64, one per lane of a wave, and no
// constraint's entries straddle two chunks: a wave adds a constraint's
// entries by a segmented sum over at most kLight lanes and the head lane
// updates acc, so each row takes one update per visit, in block order.
//
// Visits are software-pipelined with one barrier each: while the pairs of
// visit u are summed against R[u & 1], R of visit u + 1 is formed from the
// rows staged in S[(u + 1) & 1], the rows of visit u + 2 are loaded and then
// stored to S[u & 1], and the first record chunks of visit u + 1 and the
// descriptor of visit u + 3 are loaded.
constexpr int kMultiE = 4;   // entries of A_j staged (more: read from X, Z^-1 directly)
constexpr int kMultiT = 512; // threads per work-group (one Schur column)
constexpr int kMultiR = 4;   // record chunks per wave loaded one visit ahead
constexpr int kMultiP = (kMultiN * (kMultiN + 1) / 2 + kMultiT - 1) / kMultiT;  // positions per thread

// row of the upper-triangle position p = hi (hi + 1) / 2 + lo
__device__ __forceinline__ int tri_hi(int p) {
  int h = (int)((sqrtf((float)(8 * p + 1)) - 1.0f) * 0.5f);
  if (h * (h + 1) / 2 > p) --h;
  if ((h + 1) * (h + 2) / 2 <= p) ++h;
  return h;
}

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Visit records (kVisInt4 int4 each, in column order): [0] first record of
// the block, chunks, n | cj << 8 | (j takes F3) << 16 (cj = A_j's entries in
// the block), j's chunk; [1] j's NonZeroEffect, estart of A_j's entries, the
// block's packed offset (two words); [2] a | b << 8 of A_j's first two
// entries, of the next two (16 bits each); [3], [4] their values.  A ring of
// four in LDS holds the visits u .. u + 3 (written one visit ahead by five
// lanes' vector loads, read back as broadcasts).
__global__ __launch_bounds__(kMultiT) __attribute__((amdgpu_waves_per_eu(4))) void k_f3_multi(
    const int4* __restrict__ vis, const int* __restrict__ jptr, const int4* __restrict__ mrec,
    const int* __restrict__ er, const int* __restrict__ ec, const double* __restrict__ ev,
    const double* __restrict__ X, const double* __restrict__ Zi, double* __restrict__ B, int m,
    int pmax, int nmax, int soff) {
  extern __shared__ double acc[];
  __shared__ int4 ring[4][kVisInt4];
  double* R = acc + m;     // [2][pmax]
  double* S = acc + soff;  // [2][kMultiE][nmax] x {U1, W1, U2, W2}
  const int sstride = kMultiE * nmax * 4;
  const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int r = j + tid; r < m; r += kMultiT) acc[r] = 0.0;
  const int t0 = jptr[j], nv = jptr[j + 1] - t0;
  // this thread's positions p = tid + kMultiT k of R (the same for every
  // block): lo | hi << 16
  int plh[kMultiP];
#pragma unroll
  for (int k = 0; k < kMultiP; ++k) {
    const int p = tid + kMultiT * k, hi = tri_hi(p);
    plh[k] = (p - hi * (hi + 1) / 2) | (hi << 16);
  }
  // byte addresses of lanes lane + 1, + 2, + 4 (mod 64) for ds_bpermute
  const int bp1 = ((lane + 1) & 63) << 2, bp2 = ((lane + 2) & 63) << 2,
            bp4 = ((lane + 4) & 63) << 2;
  if (tid < kVisInt4 && nv > 0) ring[0][tid] = vis[(size_t)t0 * kVisInt4 + tid];
  int4 rc0[kMultiR], rc1[kMultiR];
#pragma unroll
  for (int r = 0; r < kMultiR; ++r) rc0[r] = rc1[r] = make_int4(0, 0, 0, 0);
  __syncthreads();
  for (int u = -2; u < nv; ++u) {
    // (a) loads: the record of visit u + 3, the staged rows of u + 2, the
    // first record chunks of u + 1
    int4 vq = make_int4(0, 0, 0, 0);
    if (u + 3 < nv && tid < kVisInt4) vq = vis[(size_t)(t0 + u + 3) * kVisInt4 + tid];
    int se = -1, sx = 0;
    double sv[4] = {0.0, 0.0, 0.0, 0.0}, sc = 0.0;
    if (u + 2 < nv) {
      const int4* V = ring[(u + 2) & 3];
      const int4 q0 = V[0];
      const int n2 = wave_uniform(q0.z & 255), ne = min(wave_uniform((q0.z >> 8) & 15), kMultiE);
      if (tid < ne * n2) {
        se = tid / n2;
        sx = tid - se * n2;
        const int4 q1 = V[1], q2 = V[2], qv = V[3 + (se >> 1)];
        const int ab = ((se < 2 ? q2.x : q2.y) >> (16 * (se & 1))) & 0xffff;
        const int a = ab & 255, b = ab >> 8;
        const double v = (se & 1) ? __hiloint2double(qv.w, qv.z) : __hiloint2double(qv.y, qv.x);
        sc = a != b ? v : 0.5 * v;
        const long off = (long)(((unsigned long)(unsigned)q1.w << 32) | (unsigned)q1.z);
        const double* Xq = X + off;
        const double* Zq = Zi + off;
        sv[0] = Xq[(size_t)n2 * a + sx];
        sv[1] = Zq[(size_t)n2 * b + sx];
        sv[2] = Xq[(size_t)n2 * b + sx];
        sv[3] = Zq[(size_t)n2 * a + sx];
      }
    }
    if (u + 1 >= 0 && u + 1 < nv) {
      const int4 q0 = ring[(u + 1) & 3][0];
      const int rec0 = wave_uniform(q0.x), nch = wave_uniform(q0.y), c0 = wave_uniform(q0.w);
#pragma unroll
      for (int r = 0; r < kMultiR; ++r) {
        const int c = c0 + w + 8 * r;
        rc1[r] = c < nch ? mrec[(size_t)rec0 + (size_t)c * 64 + lane] : make_int4(0, 0, 0, 0);
      }
    }
    // (b) R of visit u + 1 from S[(u + 1) & 1]
    if (u + 1 >= 0 && u + 1 < nv) {
      const int4 q0 = ring[(u + 1) & 3][0];
      const int n1 = wave_uniform(q0.z & 255), cj = wave_uniform((q0.z >> 8) & 15);
      const int P1 = n1 * (n1 + 1) / 2, ne = min(cj, kMultiE);
      const double2* S2 = reinterpret_cast<const double2*>(S + ((u + 1) & 1) * sstride);
      double* Rw = R + ((u + 1) & 1) * pmax;
#pragma unroll
      for (int k = 0; k < kMultiP; ++k) {
        const int p = tid + kMultiT * k;
        if (p >= P1) break;
        const int lo = plh[k] & 0xffff, hi = plh[k] >> 16;
        double r = 0.0;
        for (int e = 0; e < ne; ++e) {
          const double2 l0 = S2[2 * (e * nmax + lo)], l1 = S2[2 * (e * nmax + lo) + 1];
          const double2 h0 = S2[2 * (e * nmax + hi)], h1 = S2[2 * (e * nmax + hi) + 1];
          r += l0.x * h0.y + h0.x * l0.y + l1.x * h1.y + h1.x * l1.y;
        }
        if (cj > kMultiE) {
          // entries beyond the staged ones, read from X and Z^-1 directly
          const int4 q1 = ring[(u + 1) & 3][1];
          const long off = (long)(((unsigned long)(unsigned)q1.w << 32) | (unsigned)q1.z);
          const double* Xq = X + off;
          const double* Zq = Zi + off;
          for (int e = kMultiE; e < cj; ++e) {
            const int a = er[q1.y + e], b = ec[q1.y + e];
            const double v = ev[q1.y + e], c = a != b ? v : 0.5 * v;
            r += c * (Xq[(size_t)n1 * a + lo] * Zq[(size_t)n1 * b + hi] +
                      Xq[(size_t)n1 * a + hi] * Zq[(size_t)n1 * b + lo] +
                      Xq[(size_t)n1 * b + lo] * Zq[(size_t)n1 * a + hi] +
                      Xq[(size_t)n1 * b + hi] * Zq[(size_t)n1 * a + lo]);
          }
        }
        Rw[p] = lo != hi ? r : 0.5 * r;
      }
    }
    // (c) the pairs of visit u against R[u & 1]
    if (u >= 0) {
      const double* Rr = R + (u & 1) * pmax;
      const int4 q0 = ring[u & 3][0], q1 = ring[u & 3][1];
      const int rec0 = wave_uniform(q0.x), nch = wave_uniform(q0.y), c0 = wave_uniform(q0.w);
      const int fj3 = wave_uniform((q0.z >> 16) & 1), nzj = wave_uniform(q1.x);
      // kMultiR chunks of one wave at once (records past the list are
      // zeros: no head), so that their LDS reads overlap
      auto chunks = [&](const int4* rec) {
        double x[kMultiR];
#pragma unroll
        for (int r = 0; r < kMultiR; ++r)
          x[r] = __hiloint2double(rec[r].w, rec[r].z) * Rr[(rec[r].x >> 13) & 4095];
        // segmented sums over the lanes of a constraint (rem after this one)
        const int bps[3] = {bp1, bp2, bp4};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          int lo32[kMultiR], hi32[kMultiR];
#pragma unroll
          for (int r = 0; r < kMultiR; ++r) {
            lo32[r] = __builtin_amdgcn_ds_bpermute(bps[k], __double2loint(x[r]));
            hi32[r] = __builtin_amdgcn_ds_bpermute(bps[k], __double2hiint(x[r]));
          }
#pragma unroll
          for (int r = 0; r < kMultiR; ++r)
            if ((1 << k) <= ((rec[r].x >> 25) & 7)) x[r] += __hiloint2double(hi32[r], lo32[r]);
        }
#pragma unroll
        for (int r = 0; r < kMultiR; ++r) {
          const int meta = rec[r].x, crow = meta & 8191;
          const bool jo = !(rec[r].y > nzj);  // owner: tie -> the smaller index (j)
          if (((meta >> 28) & 1) && crow >= j && (jo ? fj3 : (meta >> 29) & 1))
            acc[crow] += x[r];
        }
      };
      chunks(rc0);
      for (int c1 = c0 + w + 8 * kMultiR; c1 < nch; c1 += 8 * kMultiR) {
        int4 rq[kMultiR];
#pragma unroll
        for (int r = 0; r < kMultiR; ++r) {
          const int c = c1 + 8 * r;
          rq[r] = c < nch ? mrec[(size_t)rec0 + (size_t)c * 64 + lane] : make_int4(0, 0, 0, 0);
        }
        chunks(rq);
      }
    }
    // (d) the staged rows of visit u + 2 into S[u & 1], the record of visit
    // u + 3 into the ring (the slot of visit u - 1, done)
    if (se >= 0) {
      double2* S2 = reinterpret_cast<double2*>(S + ((u + 2) & 1) * sstride);
      S2[2 * (se * nmax + sx)] = make_double2(sc * sv[0], sv[1]);
      S2[2 * (se * nmax + sx) + 1] = make_double2(sc * sv[2], sv[3]);
    }
    if (u + 3 < nv && tid < kVisInt4) ring[(u + 3) & 3][tid] = vq;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kMultiR; ++r) rc0[r] = rc1[r];
  }
  for (int r = j + tid; r < m; r += kMultiT) B[(size_t)r + (size_t)m * j] = acc[r];
}

// dynamic LDS of k_f3_multi (doubles): acc, R and S, S on a 32-byte boundary
size_t f3multi_lds(int m, int pmax, int nmax, int* soff) {
  const int so = (m + 2 * pmax + 3) & ~3;
  if (soff) *soff = so;
  return (size_t)so + (size_t)2 * kMultiE * nmax * 4;
}

int f3multi_prepare(int m, int pmax, int nmax) {
  if (m > kF3MultiMaxM || nmax > kMultiN) return -1;
  SDPA_HIP_CHECK(hipFuncSetAttribute(
      (const void*)k_f3_multi, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)(sizeof(double) * f3multi_lds(m, pmax, nmax, nullptr))));
  return 0;
}

// One thread per pair (ki >= kj) of the block's constraint list.
// grid = (ceil(ncon/64), ceil(ncon/4)), block = 256 (64 rows x 4 columns).
__global__ __launch_bounds__(256) void k_f3_light(
    const int* __restrict__ con, const int* __restrict__ estart, const int* __restrict__ ecnt,
    const int* __restrict__ nzeff, const int* __restrict__ form, const int* __restrict__ er,
    const int* __restrict__ ec, const double* __restrict__ ev, int ncon, int n,
    const double* __restrict__ X, const double* __restrict__ Zi, double* __restrict__ B, int m,
    int store) {
  if ((int)(blockIdx.x * 64 + 63) < (int)(blockIdx.y * 4)) return;
  const int ki = blockIdx.x * 64 + (threadIdx.x & 63);
  const int kj = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (ki >= ncon || kj >= ncon || ki < kj) return;
  const int nzi = nzeff[ki], nzj = nzeff[kj];
  const int o = (nzi > nzj) ? ki : kj;  // tie -> smaller index kj
  const int p = (o == ki) ? kj : ki;
  const int co = ecnt[o], cp = ecnt[p];
  double val = 0.0;
  bool mine = (form[o] == 3) && co <= kLight && cp <= kLight;
  if (mine) val = f3_pair(er, ec, ev, estart[o], co, estart[p], cp, n, X, Zi, 0, 1);
  double* dst = B + (size_t)con[ki] + (size_t)m * con[kj];
  if (store) *dst = val;
  else if (mine) *dst += val;
}

constexpr int kColsWG = 2;      // columns per work-group pass of k_f3_cols
constexpr int kColsMaxN = 256;  // block sizes whose staged rows fit LDS
constexpr int kColsRows = 4096; // rows per work-group of k_f3_cols

// Light F3 pairs by column blocks.  A work-group takes the columns
// [kc, kc + 2) of the block's constraint list (and, for balance, the mirror
// block counted from the end) and walks the rows ki >= kc, one thread per
// row.  For a column with one stored entry (a, b) the rows a and b of X and
// Z^-1 -- everything its pairs with one-entry rows read -- are staged in LDS
// first, so those pairs' eight gathers are LDS reads and the stores of B
// (column-major) are coalesced along ki.  Pairs with a multi-entry
// constraint go through f3_pair from global memory.  The value of a pair
// does not depend on which side is the owner (the calF3 terms are the same
// four products in another order); ownership decides only whether F3 or the
// F1/F2 path computes it and whether it is light.  Heavy pairs are left to
// the heavy kernels (stored as 0 when `store`).  A constraint's metadata is
// one 16-byte record (rmeta: con, estart, nzeff, packed count/form/entry
// position) and its entry's value (rval), two coalesced loads per row.
// (Interleaving the staged rows as 16-byte pairs was measured slower: more
// LDS bank conflicts on the random-row reads.)  grid = (ceil(nblk / 2),
// row segments).
__global__ __launch_bounds__(256) void k_f3_cols(
    const int4* __restrict__ rmeta, const double* __restrict__ rval, const int* __restrict__ er,
    const int* __restrict__ ec, const double* __restrict__ ev, int ncon, int n,
    const double* __restrict__ X, const double* __restrict__ Zi, double* __restrict__ B, int m,
    int store) {
  __shared__ double sR[kColsWG][4][kColsMaxN];  // X row a, X row b, Z^-1 row a, Z^-1 row b
  __shared__ int4 smeta[kColsWG];
  __shared__ double sv[kColsWG];
  const int tid = threadIdx.x;
  const int nblk = (ncon + kColsWG - 1) / kColsWG;
  // rows of the two blocks, concatenated, cut into segments of kColsRows
  // (blockIdx.y): more work-groups in flight, each with a short row walk
  int off = 0;
  const int s0 = (int)blockIdx.y * kColsRows, s1 = s0 + kColsRows;
  for (int side = 0; side < 2; ++side) {
    const int blk = side ? nblk - 1 - (int)blockIdx.x : (int)blockIdx.x;
    if (side && blk <= (int)blockIdx.x) break;
    const int kc = blk * kColsWG;
    const int len = ncon - kc;
    const int r0 = kc + max(0, s0 - off), r1 = kc + min(len, s1 - off);
    off += len;
    if (r0 >= r1) continue;
    __syncthreads();  // the previous pass is done with the staged rows
    if (tid < kColsWG) {
      const int kj = kc + tid;
      smeta[tid] = kj < ncon ? rmeta[kj] : make_int4(0, 0, 0, 0);
      sv[tid] = kj < ncon ? rval[kj] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < kColsWG * n; e += 256) {
      const int c = e / n, g = e - c * n;
      const int w = smeta[c].w;
      if ((w & 15) == 1) {
        const size_t ra = (size_t)n * ((w >> 8) & 255), rb = (size_t)n * ((w >> 16) & 255);
        sR[c][0][g] = X[ra + g];
        sR[c][1][g] = X[rb + g];
        sR[c][2][g] = Zi[ra + g];
        sR[c][3][g] = Zi[rb + g];
      }
    }
    __syncthreads();
    // the columns' metadata in registers for the whole walk
    int cnz[kColsWG], ccnt[kColsWG], cform[kColsWG], cst[kColsWG];
    size_t ccol[kColsWG];
    bool cdiag[kColsWG];
    double cv[kColsWG];
#pragma unroll
    for (int c = 0; c < kColsWG; ++c) {
      const int4 q = smeta[c];
      cnz[c] = q.z;
      ccnt[c] = q.w & 15;
      cform[c] = (q.w >> 4) & 3;
      cst[c] = q.y;
      ccol[c] = (size_t)m * q.x;
      cdiag[c] = ((q.w >> 8) & 255) == ((q.w >> 16) & 255);
      cv[c] = sv[c];
    }
    // a row's record is loaded one iteration ahead (a clamped row index stays
    // in range)
    int ki = r0 + tid;
    int kl = min(ki, ncon - 1);
    int4 qi = rmeta[kl];
    double vi = rval[kl];
    for (; ki < r1; ki += 256) {
      kl = min(ki + 256, ncon - 1);
      const int4 qn = rmeta[kl];
      const double vn = rval[kl];
      const int ci = qi.w & 15, fi = (qi.w >> 4) & 3, g = (qi.w >> 8) & 255,
                d = (qi.w >> 16) & 255;
      const size_t rowB = (size_t)qi.x;
#pragma unroll
      for (int c = 0; c < kColsWG; ++c) {
        const int kj = kc + c;
        if (kj >= ncon || ki < kj) continue;
        const int cj = ccnt[c];
        const bool jo = !(qi.z > cnz[c]);  // owner: tie -> the smaller index kj
        const int co = jo ? cj : ci, cq = jo ? ci : cj, fo = jo ? cform[c] : fi;
        const bool mine = fo == 3 && co <= kLight && cq <= kLight;
        double val = 0.0;
        if (mine) {
          if (ci == 1 && cj == 1) {
            const double* Xa = sR[c][0];
            const double* Xb = sR[c][1];
            const double* Za = sR[c][2];
            const double* Zb = sR[c][3];
            double plu;
            if (!cdiag[c]) {
              plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] + Za[d] * Xb[g] + Za[g] * Xb[d]
                             : Zb[d] * Xa[g] + Za[d] * Xb[g];
            } else {
              plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] : Zb[d] * Xa[g];
            }
            val = cv[c] * (vi * plu);
          } else {
            val = jo ? f3_pair(er, ec, ev, cst[c], cj, qi.y, ci, n, X, Zi, 0, 1)
                     : f3_pair(er, ec, ev, qi.y, ci, cst[c], cj, n, X, Zi, 0, 1);
          }
        }
        double* dst = B + rowB + ccol[c];
        if (store) *dst = val;
        else if (mine) *dst += val;
      }
      qi = qn;
      vi = vn;
    }
  }
}

// One wave per pair that involves a heavy list.  grid = (ceil(ncon/4), nheavy).
__global__ __launch_bounds__(256) void k_f3_heavy(
    const int* __restrict__ con, const int* __restrict__ estart, const int* __restrict__ ecnt,
    const int* __restrict__ nzeff, const int* __restrict__ form, const int* __restrict__ heavy,
    const int* __restrict__ er, const int* __restrict__ ec, const double* __restrict__ ev, int ncon,
    int n, const double* __restrict__ X, const double* __restrict__ Zi, double* __restrict__ B,
    int m) {
  const int h = heavy[blockIdx.y];
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= ncon) return;
  if (ecnt[k] > kLight) return;  // heavy-heavy pairs: k_f3_heavy_pair
  const int ki = k > h ? k : h, kj = k > h ? h : k;
  const int nzi = nzeff[ki], nzj = nzeff[kj];
  const int o = (nzi > nzj) ? ki : kj;
  const int p = (o == ki) ? kj : ki;
  if (form[o] != 3) return;
  double v = f3_pair<true>(er, ec, ev, estart[o], ecnt[o], estart[p], ecnt[p], n, X, Zi, lane, 64);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0) B[(size_t)con[ki] + (size_t)m * con[kj]] += v;
}

// Pairs of two heavy lists: kHeavySlices work-groups per pair, each taking
// every kHeavySlices-th stripe of 256 of the owner x partner term grid, four
// products per thread per round with their loads issued together; the
// slices' sums go to `part` and k_f3_heavy_sum adds them to B in slice
// order (deterministic).  grid = (nheavy, nheavy, kHeavySlices).
constexpr int kHeavySlices = 16;

__global__ __launch_bounds__(256) void k_f3_heavy_pair(
    const int* __restrict__ estart, const int* __restrict__ ecnt, const int* __restrict__ nzeff,
    const int* __restrict__ form, const int* __restrict__ heavy, int nheavy,
    const int* __restrict__ er, const int* __restrict__ ec, const double* __restrict__ ev, int n,
    const double* __restrict__ X, const double* __restrict__ Zi, double* __restrict__ part) {
  const int h1 = heavy[blockIdx.x], h2 = heavy[blockIdx.y];
  if (h1 < h2) return;
  const int ki = h1, kj = h2;
  const int o = (nzeff[ki] > nzeff[kj]) ? ki : kj;
  const int p = (o == ki) ? kj : ki;
  if (form[o] != 3) return;
  const int so = estart[o], co = ecnt[o], sp = estart[p], cp = ecnt[p];
  const long total = (long)co * cp;
  constexpr int kStep = 256 * kHeavySlices;
  // (e1, e2) of t = e1 cp + e2, advanced by kStep without a 64-bit division
  const int qs = kStep / cp, rs = kStep % cp;
  const int t00 = 256 * (int)blockIdx.z + (int)threadIdx.x;
  int e1c = t00 / cp, e2c = t00 % cp;
  double acc = 0.0;
  for (long t0 = t00; t0 < total; t0 += 4L * kStep) {
    int a[4], b[4], g[4], d[4];
    double w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = t0 + (long)kStep * u < total;
      const int e1 = in ? e1c : 0, e2 = in ? e2c : 0;
      e1c += qs;
      e2c += rs;
      if (e2c >= cp) {
        e2c -= cp;
        ++e1c;
      }
      a[u] = er[so + e1];
      b[u] = ec[so + e1];
      g[u] = er[sp + e2];
      d[u] = ec[sp + e2];
      w[u] = in ? ev[so + e1] * ev[sp + e2] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double* Xa = X + (size_t)n * a[u];
      const double* Xb = X + (size_t)n * b[u];
      const double* Za = Zi + (size_t)n * a[u];
      const double* Zb = Zi + (size_t)n * b[u];
      const int gg = g[u], dd = d[u];
      double plu;
      if (a[u] != b[u]) {
        if (gg != dd) plu = Zb[dd] * Xa[gg] + Zb[gg] * Xa[dd] + Za[dd] * Xb[gg] + Za[gg] * Xb[dd];
        else plu = Zb[dd] * Xa[gg] + Za[dd] * Xb[gg];
      } else {
        if (gg != dd) plu = Zb[dd] * Xa[gg] + Zb[gg] * Xa[dd];
        else plu = Zb[dd] * Xa[gg];
      }
      acc += w[u] * plu;
    }
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    part[((size_t)blockIdx.x * nheavy + blockIdx.y) * kHeavySlices + blockIdx.z] =
        (red[0] + red[1]) + (red[2] + red[3]);
}

// B[i, j] += the slices of each heavy pair, in slice order.  One thread per
// (x, y) of the nheavy x nheavy grid.
__global__ __launch_bounds__(256) void k_f3_heavy_sum(
    const int* __restrict__ con, const int* __restrict__ nzeff, const int* __restrict__ form,
    const int* __restrict__ heavy, int nheavy, const double* __restrict__ part,
    double* __restrict__ B, int m) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nheavy * nheavy) return;
  const int x = t / nheavy, y = t % nheavy;
  const int h1 = heavy[x], h2 = heavy[y];
  if (h1 < h2) return;
  const int o = (nzeff[h1] > nzeff[h2]) ? h1 : h2;
  if (form[o] != 3) return;
  const double* ps = part + (size_t)t * kHeavySlices;
  double sum = 0.0;
  for (int z = 0; z < kHeavySlices; ++z) sum += ps[z];
  B[(size_t)con[h1] + (size_t)m * con[h2]] += sum;
}

// ---- F1/F2 ----------------------------------------------------------------
// Dense copy of the owners' A_j (sparse owners are expanded, dense ones copied
// from the pool) for one chunk.  grid = (chunk), block = 256.
// T is the element type of the products (double, or float for the FP32
// assembly); the pool is in the same type.
template <class T>
__global__ __launch_bounds__(256) void k_densify(const int* __restrict__ own, int first,
                                                 const int* __restrict__ estart,
                                                 const int* __restrict__ ecnt,
                                                 const int* __restrict__ dpos,
                                                 const int* __restrict__ er,
                                                 const int* __restrict__ ec,
                                                 const double* __restrict__ ev,
                                                 const T* __restrict__ pool, int n,
                                                 T* __restrict__ out) {
  const int j = own[first + blockIdx.x];
  T* A = out + (size_t)blockIdx.x * n * n;
  const size_t nn = (size_t)n * n;
  if (dpos[j] >= 0) {
    const T* src = pool + (size_t)dpos[j] * nn;
    for (size_t e = threadIdx.x; e < nn; e += 256) A[e] = src[e];
    return;
  }
  for (size_t e = threadIdx.x; e < nn; e += 256) A[e] = T(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int e = 0; e < ecnt[j]; ++e) {
      int r = er[estart[j] + e], c = ec[estart[j] + e];
      A[r + (size_t)n * c] += (T)ev[estart[j] + e];
      if (r != c) A[c + (size_t)n * r] += (T)ev[estart[j] + e];
    }
  }
}

// C[i] = sum over s of P[s * len + i], in order of s (in FP64)
template <class T>
__global__ __launch_bounds__(256) void k_sum_slices(const T* __restrict__ P, int ns, size_t len,
                                                    double* __restrict__ C) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  double a = 0.0;
  for (int s = 0; s < ns; ++s) a += (double)P[(size_t)s * len + i];
  C[i] = a;
}

// dst = float(src), n elements
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ src,
                                                float* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    dst[i] = (float)src[i];
}

__device__ __forceinline__ bool owns(int jloc, int iloc, const int* nzeff) {
  const int jnz = nzeff[jloc], inz = nzeff[iloc];
  return !((jnz < inz) || (inz == jnz && iloc < jloc));
}

// Sparse partners of F1 owners: one wave per (owner in chunk, partner).
// grid = (ceil(ncon/4), chunk).
template <class T>
__global__ __launch_bounds__(256) void k_f1_sparse(
    const int* __restrict__ own, int first, const int* __restrict__ con,
    const int* __restrict__ estart, const int* __restrict__ ecnt, const int* __restrict__ nzeff,
    const int* __restrict__ dpos, const int* __restrict__ er, const int* __restrict__ ec,
    const double* __restrict__ ev, int ncon, int n, const T* __restrict__ G,
    double* __restrict__ B, int m) {
  const int jc = blockIdx.y;
  const int j = own[first + jc];
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= ncon || dpos[i] >= 0 || !owns(j, i, nzeff)) return;
  const T* Gj = G + (size_t)jc * n * n;
  double v = 0.0;
  for (int e = lane; e < ecnt[i]; e += 64) {
    int r = er[estart[i] + e], c = ec[estart[i] + e];
    const double g = (r == c) ? (double)Gj[r + (size_t)n * c]
                              : (double)Gj[r + (size_t)n * c] + (double)Gj[c + (size_t)n * r];
    v += ev[estart[i] + e] * g;
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0) {
    int gi = con[i], gj = con[j];
    int a = gi > gj ? gi : gj, b = gi > gj ? gj : gi;
    B[(size_t)a + (size_t)m * b] += v;
  }
}

// Dense partners: P = Abar^T Gbar over the pool columns rlo..ndense-1
// ((ndense - rlo) x chunk) was formed by a GEMM; add the owned entries.
// grid = (ceil((ndense - rlo)/256), chunk).
__global__ __launch_bounds__(256) void k_f1_scatter(const int* __restrict__ own, int first,
                                                    const int* __restrict__ dlist,
                                                    const int* __restrict__ con,
                                                    const int* __restrict__ nzeff, int ndense,
                                                    int rlo, const double* __restrict__ P,
                                                    double* __restrict__ B, int m) {
  const int jc = blockIdx.y;
  const int id = rlo + blockIdx.x * 256 + threadIdx.x;
  if (id >= ndense) return;
  const int j = own[first + jc];
  const int i = dlist[id];
  if (!owns(j, i, nzeff)) return;
  int gi = con[i], gj = con[j];
  int a = gi > gj ? gi : gj, b = gi > gj ? gj : gi;
  B[(size_t)a + (size_t)m * b] += P[(id - rlo) + (size_t)(ndense - rlo) * jc];
}

// The first dense-pool column that an owner of the chunk [first, first + cnt)
// of T's F1 owners owns (ndense if none): the rule of owns() on the host.
static int f1_first_owned(const BlockTable& T, int first, int cnt) {
  std::vector<int> local(T.ndense, -1);
  for (int k = 0; k < T.ncon; ++k)
    if (T.h_dpos[k] >= 0) local[T.h_dpos[k]] = k;
  for (int id = 0; id < T.ndense; ++id) {
    const int i = local[id];
    for (int t = first; t < first + cnt; ++t) {
      const int j = T.h_f1own[t];
      const int jnz = T.h_nzeff[j], inz = T.h_nzeff[i];
      if (!((jnz < inz) || (inz == jnz && i < j))) return id;
    }
  }
  return T.ndense;
}

// ---- LP ---------------------------------------------------------------------
// One thread per LP coordinate; pairs k1 <= k2 of its constraint list.
__global__ void k_lp_bmat(const int* __restrict__ ptr, const int* __restrict__ lcon,
                          const double* __restrict__ lval, int nLP, const double* __restrict__ X,
                          const double* __restrict__ Zi, long lpOff, double* __restrict__ B,
                          int m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nLP) return;
  const double xz = X[lpOff + t] * Zi[lpOff + t];
  for (int k1 = ptr[t]; k1 < ptr[t + 1]; ++k1) {
    const int j = lcon[k1];
    const double aj = lval[k1];
    for (int k2 = k1; k2 < ptr[t + 1]; ++k2) {
      const int i = lcon[k2];
      const double v = xz * lval[k2] * aj;
      const int a = i > j ? i : j, b = i > j ? j : i;
      atomicAdd(B + (size_t)a + (size_t)m * b, v);
    }
  }
}

struct F1Scratch;

// The products of f1_block in the element type T.
static int gemm_t(hipStream_t s, int ta, int tb, int M, int N, int K, const double* A, int lda,
                  long sA, const double* B, int ldb, long sB, double* C, int ldc, long sC,
                  int batch) {
  return dgemm(s, ta, tb, M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, C, ldc, sC, batch);
}
static int gemm_t(hipStream_t s, int ta, int tb, int M, int N, int K, const float* A, int lda,
                  long sA, const float* B, int ldb, long sB, float* C, int ldc, long sC,
                  int batch) {
  return sgemm(s, ta, tb, M, N, K, 1.0f, A, lda, sA, B, ldb, sB, 0.0f, C, ldc, sC, batch);
}

// F1/F2 owners of one block: G_j = X A_j Z^-1 for chunks of owners, then
// A_i . G_j for their partners.  T = double: FP64 throughout.  T = float
// (the split-precision assembly): X, Z^-1 and the dense pool are rounded to
// FP32 and the three products run on FP32 MFMA with FP32 accumulation; the
// slices of the long Abar^T Gbar product are summed in FP64, the sparse
// partners' gathers are FP64 sums of FP32 values, and B stays FP64.
template <class T>
static int f1_block(Ctx* c, BlockTable& T_, const double* X64, const double* Zi64) {
  BlockTable& Tb = T_;
  const int n = Tb.n;
  const size_t nn = (size_t)n * n;
  hipStream_t s = c->stream;
  // chunk so that densified A, X*A and G each fit in a 3 x 512 MiB budget
  // (in FP64; the FP32 assembly takes the same chunks)
  size_t budget = (size_t)512 << 20;
  int chunk = (int)std::max<size_t>(1, budget / (nn * sizeof(double)));
  // whole 128-column tiles of the Abar^T Gbar product when the chunk is wide
  if (chunk >= 128) chunk -= chunk % 128;
  if (chunk > Tb.nF1) chunk = Tb.nF1;
  size_t need = 3 * (size_t)chunk * nn;
  if (need > c->capG) {
    if (c->d_G) SDPA_HIP_CHECK(hipFree(c->d_G));
    SDPA_HIP_CHECK(hipMalloc(&c->d_G, need * sizeof(double)));
    c->capG = need;
  }
  T* Ad = reinterpret_cast<T*>(c->d_G);
  T* XA = Ad + (size_t)chunk * nn;
  T* G = Ad + 2 * (size_t)chunk * nn;
  if (Tb.ndense > 0) {
    size_t needP = (size_t)Tb.ndense * chunk;
    if (needP > c->capF1C) {
      if (c->d_F1C) SDPA_HIP_CHECK(hipFree(c->d_F1C));
      SDPA_HIP_CHECK(hipMalloc(&c->d_F1C, needP * sizeof(double)));
      c->capF1C = needP;
    }
  }
  // the operands in T
  const T* X;
  const T* Zi;
  const T* pool;
  if constexpr (sizeof(T) == 8) {
    X = X64;
    Zi = Zi64;
    pool = Tb.d_dense;
  } else {
    if (2 * nn > c->capXZ32) {
      if (c->d_xz32) SDPA_HIP_CHECK(hipFree(c->d_xz32));
      SDPA_HIP_CHECK(hipMalloc(&c->d_xz32, 2 * nn * sizeof(float)));
      c->capXZ32 = 2 * nn;
    }
    const unsigned gb = (unsigned)std::min<size_t>(1024, (nn + 255) / 256);
    hipLaunchKernelGGL(k_to_f32, dim3(gb), dim3(256), 0, s, X64, c->d_xz32, nn);
    hipLaunchKernelGGL(k_to_f32, dim3(gb), dim3(256), 0, s, Zi64, c->d_xz32 + nn, nn);
    X = c->d_xz32;
    Zi = c->d_xz32 + nn;
    if (Tb.ndense > 0 && !Tb.d_dense32) {
      // the pool in FP32, made at the block's first FP32 assembly
      const size_t len = nn * Tb.ndense;
      SDPA_HIP_CHECK(hipMalloc(&Tb.d_dense32, len * sizeof(float)));
      hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::min<size_t>(8192, (len + 255) / 256)),
                         dim3(256), 0, s, Tb.d_dense, Tb.d_dense32, len);
    }
    pool = Tb.d_dense32;
  }
  // dense-member list (local indices of dense constraints, in pool order),
  // built at load time
  const int* d_dlist = Tb.d_dlist;
  if (Tb.ndense > 0 && (Tb.f1chunk != chunk || Tb.f1rlo.empty())) {
    Tb.f1chunk = chunk;
    Tb.f1rlo.clear();
    for (int first = 0; first < Tb.nF1; first += chunk)
      Tb.f1rlo.push_back(f1_first_owned(Tb, first, std::min(chunk, Tb.nF1 - first)));
  }
  for (int first = 0; first < Tb.nF1; first += chunk) {
    const int cnt = std::min(chunk, Tb.nF1 - first);
    // the chunk's A_t: straight from the dense pool when its owners are
    // consecutive pool columns, else densified into Ad
    const T* At = Ad;
    {
      const int p0 = Tb.h_dpos[Tb.h_f1own[first]];
      bool consecutive = p0 >= 0;
      for (int t = 1; t < cnt && consecutive; ++t)
        consecutive = Tb.h_dpos[Tb.h_f1own[first + t]] == p0 + t;
      if (consecutive)
        At = pool + (size_t)p0 * nn;
      else
        hipLaunchKernelGGL(k_densify<T>, dim3(cnt), dim3(256), 0, s, Tb.d_f1own, first,
                           Tb.d_estart, Tb.d_ecnt, Tb.d_dpos, c->d_er, c->d_ec, c->d_ev, pool, n,
                           Ad);
    }
    // XA_t = X A_t and G_t = XA_t Z^-1 for the chunk (strided-batched MFMA GEMMs)
    if (int e = gemm_t(s, 0, 0, n, n, n, X, n, 0, At, n, (long)nn, XA, n, (long)nn, cnt))
      return e;
    if (int e = gemm_t(s, 0, 0, n, n, n, XA, n, (long)nn, Zi, n, 0, G, n, (long)nn, cnt))
      return e;
    const int rlo = Tb.ndense > 0 ? Tb.f1rlo[first / chunk] : 0;
    if (Tb.ndense > rlo) {
      // P = Abar^T * Gbar : (ndense x n^2) * (n^2 x cnt), over the pool
      // columns from rlo on (the ones before are owned by no owner of the
      // chunk: for dense constraints of equal size the pair belongs to the
      // smaller index, so a chunk needs only the triangle below it).  The
      // output has few tiles for its long inner dimension: the n^2 rows are
      // split into ks slices (one batched call, one partial product each) and
      // the partials summed in FP64, so that the GEMM fills the chip.
      const T* Abar = pool + (size_t)rlo * nn;
      const int M = Tb.ndense - rlo, N = cnt;
      const long K = (long)nn;
      const long tiles = (long)((M + 127) / 128) * ((N + 127) / 128);
      int ks = (int)std::min<long>(32, std::max<long>(1, 1024 / tiles));
      if (K / ks < 1024) ks = (int)std::max<long>(1, K / 1024);
      const size_t len = (size_t)M * N;
      if (ks <= 1 && sizeof(T) == 8) {
        if (int e = gemm_t(s, 1, 0, M, N, (int)K, Abar, (int)nn, 0, G, (int)nn, 0,
                           reinterpret_cast<T*>(c->d_F1C), M, 0, 1))
          return e;
      } else {
        const long Kc = K / ks, rem = K - Kc * ks;
        const int nparts = ks + (rem > 0 ? 1 : 0);
        if ((size_t)nparts * len > c->capF1P) {
          if (c->d_F1P) SDPA_HIP_CHECK(hipFree(c->d_F1P));
          c->capF1P = (size_t)nparts * len;
          SDPA_HIP_CHECK(hipMalloc(&c->d_F1P, c->capF1P * sizeof(double)));
        }
        T* P = reinterpret_cast<T*>(c->d_F1P);
        if (int e = gemm_t(s, 1, 0, M, N, (int)Kc, Abar, (int)nn, Kc, G, (int)nn, Kc, P, M,
                           (long)len, ks))
          return e;
        if (rem > 0)
          if (int e = gemm_t(s, 1, 0, M, N, (int)rem, Abar + Kc * ks, (int)nn, 0, G + Kc * ks,
                             (int)nn, 0, P + (size_t)ks * len, M, 0, 1))
            return e;
        hipLaunchKernelGGL(k_sum_slices<T>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                           P, nparts, len, c->d_F1C);
      }
      hipLaunchKernelGGL(k_f1_scatter, dim3((M + 255) / 256, cnt), dim3(256), 0, s, Tb.d_f1own,
                         first, d_dlist, Tb.d_con, Tb.d_nzeff, Tb.ndense, rlo, c->d_F1C, c->d_B,
                         c->m);
    }
    if (Tb.ndense < Tb.ncon)
      hipLaunchKernelGGL(k_f1_sparse<T>, dim3((Tb.ncon + 3) / 4, cnt), dim3(256), 0, s,
                         Tb.d_f1own, first, Tb.d_con, Tb.d_estart, Tb.d_ecnt, Tb.d_nzeff,
                         Tb.d_dpos, c->d_er, c->d_ec, c->d_ev, Tb.ncon, n, G, c->d_B, c->m);
    // the next chunk rewrites the descriptor slots and scratch: stream order keeps it safe
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

int make_bmat(Ctx* c, int precision) {
  hipStream_t s = c->stream;
  const int m = c->m;
  if (c->f3multi) {
    // stores every entry of the lower triangle: the first writer of B
    int soff = 0;
    const size_t lds = f3multi_lds(m, c->f3m_pmax, c->f3m_nmax, &soff);
    hipLaunchKernelGGL(k_f3_multi, dim3(m), dim3(kMultiT), sizeof(double) * lds, s,
                       (const int4*)c->d_jent, c->d_jptr, (const int4*)c->d_mrec, c->d_er,
                       c->d_ec, c->d_ev, c->d_X, c->d_Zinv, c->d_B, m, c->f3m_pmax,
                       c->f3m_nmax, soff);
  } else if (!c->bmatStoreFirst) {
    SDPA_HIP_CHECK(hipMemsetAsync(c->d_B, 0, sizeof(double) * (size_t)m * m, s));
  }
  for (int q = 0; q < c->nSDP; ++q) {
    BlockTable& T = c->blk[q];
    if (T.ncon == 0) continue;
    const double* X = c->d_X + T.off;
    const double* Zi = c->d_Zinv + T.off;
    if (T.nF1 > 0) {
      if (c->timing) SDPA_HIP_CHECK(hipEventRecord(c->ev[2], s));
      if (int e = precision == 1 ? f1_block<float>(c, T, X, Zi) : f1_block<double>(c, T, X, Zi))
        return e;
      if (c->timing) {
        // the F1/F2 share of the assembly (timing synchronizes)
        SDPA_HIP_CHECK(hipEventRecord(c->ev[3], s));
        SDPA_HIP_CHECK(hipEventSynchronize(c->ev[3]));
        float ms = 0.f;
        SDPA_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        c->t_f1 += ms;
      }
    }
    if (T.nF3 > 0) {
      if (c->f3multi && T.n <= kMultiN) {
        // light pairs done by k_f3_multi
      } else if (T.f3cols && T.n <= kColsMaxN) {
        const int nblk = (T.ncon + kColsWG - 1) / kColsWG;
        const dim3 gc((nblk + 1) / 2, (T.ncon + kColsWG + kColsRows - 1) / kColsRows);
        hipLaunchKernelGGL(k_f3_cols, gc, dim3(256), 0, s, (const int4*)T.d_rmeta, T.d_rval,
                           c->d_er, c->d_ec, c->d_ev, T.ncon, T.n, X, Zi, c->d_B, m,
                           c->bmatStoreFirst ? 1 : 0);
      } else {
        dim3 g((T.ncon + 63) / 64, (T.ncon + 3) / 4);
        hipLaunchKernelGGL(k_f3_light, g, dim3(256), 0, s, T.d_con, T.d_estart, T.d_ecnt,
                           T.d_nzeff, T.d_form, c->d_er, c->d_ec, c->d_ev, T.ncon, T.n, X, Zi,
                           c->d_B, m, c->bmatStoreFirst ? 1 : 0);
      }
      if (T.nheavy > 0) {
        dim3 gh((T.ncon + 3) / 4, T.nheavy);
        hipLaunchKernelGGL(k_f3_heavy, gh, dim3(256), 0, s, T.d_con, T.d_estart, T.d_ecnt,
                           T.d_nzeff, T.d_form, T.d_heavy, c->d_er, c->d_ec, c->d_ev, T.ncon, T.n,
                           X, Zi, c->d_B, m);
        const size_t np = (size_t)T.nheavy * T.nheavy * kHeavySlices;
        if (np > c->capHpart) {
          if (c->d_hpart) SDPA_HIP_CHECK(hipFree(c->d_hpart));
          c->capHpart = np;
          SDPA_HIP_CHECK(hipMalloc(&c->d_hpart, np * sizeof(double)));
        }
        hipLaunchKernelGGL(k_f3_heavy_pair, dim3(T.nheavy, T.nheavy, kHeavySlices), dim3(256), 0,
                           s, T.d_estart, T.d_ecnt, T.d_nzeff, T.d_form, T.d_heavy, T.nheavy,
                           c->d_er, c->d_ec, c->d_ev, T.n, X, Zi, c->d_hpart);
        hipLaunchKernelGGL(k_f3_heavy_sum, dim3((T.nheavy * T.nheavy + 255) / 256), dim3(256), 0,
                           s, T.d_con, T.d_nzeff, T.d_form, T.d_heavy, T.nheavy, c->d_hpart,
                           c->d_B, m);
      }
    }
  }
  if (c->nLP > 0)
    hipLaunchKernelGGL(k_lp_bmat, dim3((c->nLP + 255) / 256), dim3(256), 0, s, c->d_lpptr,
                       c->d_lpcon, c->d_lpval, c->nLP, c->d_X, c->d_Zinv, c->lpOff, c->d_B, m);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdpa_amd
