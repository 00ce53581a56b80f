// sdpa_gpu.hip -- implementation of the C ABI in include/sdpa_gpu.h:
// context lifetime, problem upload (with SDPA's structural preprocessing),
// and the thin wrappers around the kernels in dense/schur/ops/eig.hip.
#include "../../include/sdpa_gpu.h"
#include "ctx.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>

using namespace sdpa_amd;

struct sdpa_gpu_ctx {
  Ctx c;
};

namespace {

constexpr int kLightHost = 8;  // must match schur.hip's kLight

int fail(Ctx* c, int code, const char* msg) {
  c->last_error = msg;
  return code;
}

template <class T>
int upload_raw(Ctx* c, T** dst, const T* src, size_t count) {
  (void)c;
  if (*dst) SDPA_HIP_CHECK(hipFree(*dst));
  *dst = nullptr;
  SDPA_HIP_CHECK(hipMalloc((void**)dst, sizeof(T) * (count ? count : 1)));
  if (count) SDPA_HIP_CHECK(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
  return 0;
}

template <class T>
int upload(Ctx* c, T** dst, const std::vector<T>& src) {
  return upload_raw(c, dst, src.data(), src.size());
}

// Host threads for the loader's passes over the input records: the
// machine's hardware threads, at most 16, about one per 2^20 units of work.
int loader_threads(long work) {
  const unsigned hw = std::thread::hardware_concurrency();
  const long nt = std::min<long>(hw ? (long)hw : 1L, 16L);
  return (int)std::max<long>(1L, std::min<long>(nt, work >> 20));
}

// f(t, begin, end) for t = 0..nt-1 over contiguous, ordered chunks of [0, n),
// one host thread each (a chunk whose thread cannot be started runs here)
template <class F>
void par_chunks(int nt, long n, const F& f) {
  if (nt <= 1) {
    f(0, 0L, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t) {
    const long b = n * t / nt, e = n * (t + 1) / nt;
    try {
      th.emplace_back([&f, t, b, e] { f(t, b, e); });
    } catch (...) {
      f(t, b, e);
    }
  }
  for (auto& x : th) x.join();
}

// One column of the dense pool: A_k of the block (n x n, both triangles) from
// its entries (SparseMatrix::changeToDense, sdpa_struct.cpp:510-555: a later
// duplicate entry overwrites an earlier one).
void fill_dense_column(double* A, int k, int n, const std::vector<int>& estart,
                       const std::vector<int>& ecnt, const int* er, const int* ec,
                       const double* ev) {
  std::fill(A, A + (size_t)n * n, 0.0);
  for (int e = estart[k]; e < estart[k] + ecnt[k]; ++e) {
    const int r = er[e], cc = ec[e];
    A[r + (size_t)n * cc] = ev[e];
    if (r != cc) A[cc + (size_t)n * r] = ev[e];
  }
}

constexpr int kStageBufs = 4;  // at most the size of Ctx::h_stage
constexpr size_t kStageBytes = (size_t)32 << 20;

// the pinned staging ring (allocated once per context)
int ensure_stage(Ctx* c) {
  if (c->h_stage[0]) return 0;
  for (int b = 0; b < kStageBufs; ++b) {
    SDPA_HIP_CHECK(hipHostMalloc(&c->h_stage[b], kStageBytes, hipHostMallocDefault));
    SDPA_HIP_CHECK(hipEventCreateWithFlags(&c->ev_stage[b], hipEventDisableTiming));
    SDPA_HIP_CHECK(hipEventRecord(c->ev_stage[b], c->stream));
  }
  return 0;
}

int dalloc(double** p, size_t n) {
  if (*p) SDPA_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  SDPA_HIP_CHECK(hipMalloc((void**)p, sizeof(double) * (n ? n : 1)));
  SDPA_HIP_CHECK(hipMemset(*p, 0, sizeof(double) * (n ? n : 1)));
  return 0;
}

// Many small device arrays in one allocation and one copy: add() places each
// array at a 256-byte aligned offset of a host image, commit() uploads the
// image and points every added pointer at its array.
struct Arena {
  std::vector<unsigned char> h;
  std::vector<std::pair<void**, size_t>> fix;
  template <class T>
  void add(T** dst, const std::vector<T>& v) {
    const size_t off = (h.size() + 255) & ~(size_t)255;
    h.resize(off + sizeof(T) * (v.empty() ? 1 : v.size()), 0);
    if (!v.empty()) std::memcpy(h.data() + off, v.data(), sizeof(T) * v.size());
    fix.push_back({(void**)dst, off});
  }
  int commit(void** base) {
    if (*base) SDPA_HIP_CHECK(hipFree(*base));
    *base = nullptr;
    SDPA_HIP_CHECK(hipMalloc(base, h.empty() ? 1 : h.size()));
    if (!h.empty()) SDPA_HIP_CHECK(hipMemcpy(*base, h.data(), h.size(), hipMemcpyHostToDevice));
    for (auto& f : fix) *f.first = (unsigned char*)*base + f.second;
    return 0;
  }
};

void free_problem(Ctx* c) {
  sparse_schur_free(&c->ts);
  c->schurSparse = 0;
  for (auto& T : c->blk) {
    if (T.d_dense) (void)hipFree(T.d_dense);
    if (T.d_dense32) (void)hipFree(T.d_dense32);
  }
  c->blk.clear();
  if (c->d_blkarena) (void)hipFree(c->d_blkarena);
  c->d_blkarena = nullptr;
  void** ptrs[] = {(void**)&c->d_er, (void**)&c->d_ec, (void**)&c->d_ev, (void**)&c->d_lpptr,
                   (void**)&c->d_lpcon, (void**)&c->d_lpval, (void**)&c->d_cptr, (void**)&c->d_cp1,
                   (void**)&c->d_cp2, (void**)&c->d_cv, (void**)&c->d_ppos, (void**)&c->d_pptr,
                   (void**)&c->d_pcon, (void**)&c->d_pval, (void**)&c->d_b, (void**)&c->d_C,
                   (void**)&c->d_X, (void**)&c->d_Z, (void**)&c->d_y, (void**)&c->d_invCholX,
                   (void**)&c->d_invCholZ, (void**)&c->d_Zinv, (void**)&c->d_LX, (void**)&c->d_LZ,
                   (void**)&c->d_Dx, (void**)&c->d_Dz, (void**)&c->d_dy, (void**)&c->d_rzinv,
                   (void**)&c->d_xrdzinv, (void**)&c->d_g, (void**)&c->d_w1, (void**)&c->d_w2,
                   (void**)&c->d_w3, (void**)&c->d_pvec, (void**)&c->d_dmat, (void**)&c->d_X0,
                   (void**)&c->d_jptr, (void**)&c->d_jent, (void**)&c->d_mrec,
                   (void**)&c->d_Z0, (void**)&c->d_B, (void**)&c->d_Binv, (void**)&c->d_Winv, (void**)&c->d_G, (void**)&c->d_F1C, (void**)&c->d_F1P, (void**)&c->d_hpart, (void**)&c->d_xz32,
                   (void**)&c->d_info, (void**)&c->d_red, (void**)&c->d_eig,
                   (void**)&c->d_sdpN, (void**)&c->d_sdpOff};
  for (void** p : ptrs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->capG = c->capF1C = c->capF1P = c->capHpart = c->capXZ32 = 0;
  free_mixed(c);
  c->loaded = false;
}

__global__ void k_shift_diag(double* B, int m, double shift) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < m) B[i + (size_t)m * i] += shift;
}

}  // namespace

extern "C" {

int sdpa_gpu_create(int device, sdpa_gpu_ctx** out) {
  if (!out) return -1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -2;
  if (device < 0 || device >= ndev) return -3;
  SDPA_HIP_CHECK(hipSetDevice(device));
  sdpa_gpu_ctx* h = new sdpa_gpu_ctx();
  Ctx* c = &h->c;
  c->device = device;
  SDPA_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) SDPA_HIP_CHECK(hipEventCreate(&e));
  SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_red, sizeof(double) * 64, hipHostMallocDefault));
  SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_info, sizeof(int) * 64, hipHostMallocDefault));
  *out = h;
  return 0;
}

int sdpa_gpu_destroy(sdpa_gpu_ctx* h) {
  if (!h) return 0;
  Ctx* c = &h->c;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_problem(c);
  c->dw.release();
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  if (c->h_red) (void)hipHostFree(c->h_red);
  if (c->h_eig) (void)hipHostFree(c->h_eig);
  for (int b = 0; b < kStageBufs; ++b) {
    if (c->h_stage[b]) (void)hipHostFree(c->h_stage[b]);
    if (c->ev_stage[b]) (void)hipEventDestroy(c->ev_stage[b]);
  }
  if (c->h_info) (void)hipHostFree(c->h_info);
  (void)hipStreamDestroy(c->stream);
  delete h;
  return 0;
}

const char* sdpa_gpu_last_error(sdpa_gpu_ctx* h) { return h ? h->c.last_error.c_str() : "no context"; }

int sdpa_gpu_load_problem(sdpa_gpu_ctx* h, int m, int nBlock, const int* blockStruct,
                          const double* cvec, long nnz, const int* K, const int* L, const int* I,
                          const int* J, const double* V, double kappa) {
  if (!h) return -1;
  Ctx* c = &h->c;
  SDPA_HIP_CHECK(hipSetDevice(c->device));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  free_problem(c);
  if (int e = c->dw.clear_err(c->stream)) return e;
  if (m <= 0 || nBlock <= 0) return fail(c, -10, "m and nBlock must be positive");
  c->m = m;
  c->nBlock = nBlock;
  // BlockStruct::makeInternalStructure (sdpa_block.cpp:66-108)
  c->blockStruct.assign(blockStruct, blockStruct + nBlock);
  c->blockType.assign(nBlock, 0);
  c->blockNumber.assign(nBlock, 0);
  c->nSDP = 0;
  c->nLP = 0;
  c->sdpN.clear();
  for (int l = 0; l < nBlock; ++l) {
    int s = blockStruct[l];
    if (s >= 2) {
      c->blockType[l] = 1;
      c->blockNumber[l] = c->nSDP++;
      c->sdpN.push_back(s);
    } else if (s < 0 || s == 1) {
      c->blockType[l] = 2;
      c->blockNumber[l] = c->nLP;
      c->nLP += (s < 0 ? -s : 1);
    } else {
      return fail(c, -11, "block size 0");
    }
  }
  c->sdpOff.assign(c->nSDP, 0);
  long off = 0;
  int nd = 0;
  for (int q = 0; q < c->nSDP; ++q) {
    c->sdpOff[q] = off;
    off += (long)c->sdpN[q] * c->sdpN[q];
    nd += c->sdpN[q];
  }
  c->lpOff = off;
  c->total = off + c->nLP;
  c->nDim = nd + c->nLP;

  // Group the records.  The SDP entries of constraints 1..m go to buckets
  // (SDP block q, constraint k), laid out block by block and constraint by
  // constraint, each bucket's entries in input order: a stable counting sort
  // over ordered chunks of the records, one host thread per chunk, which
  // also validates them.  C = -F0 and the LP records are taken afterwards in
  // input order (duplicate LP coordinates of a constraint are summed).
  const int nS = c->nSDP;
  std::vector<double> hC(c->total, 0.0);
  std::vector<std::vector<std::pair<int, double>>> lpent(m);  // (coordinate, value)
  const long nb = (long)nS * m;
  int nt = loader_threads(nnz);
  while (nt > 1 && (long)nt * nb > (32L << 20)) --nt;  // two counters per bucket per thread
  std::vector<std::vector<int>> tcnt(nt), tdg(nt);
  std::vector<std::vector<long>> tother(nt);  // F0 and LP records of each chunk
  std::vector<long> terr(nt, -1);
  std::vector<int> tcode(nt, 0);
  {
    const int* bType = c->blockType.data();
    const int* bNum = c->blockNumber.data();
    const int* sN = c->sdpN.data();
    par_chunks(nt, nnz, [&](int t, long b0, long b1) {
      std::vector<int>& cnt = tcnt[t];
      std::vector<int>& dg = tdg[t];
      cnt.assign(nb, 0);
      dg.assign(nb, 0);
      for (long e = b0; e < b1; ++e) {
        const int k = K[e], l = L[e] - 1;
        int code = 0;
        if (k < 0 || k > m || l < 0 || l >= nBlock) {
          code = -12;
        } else if (bType[l] == 1) {
          const int q = bNum[l], n = sN[q];
          const int i = I[e] - 1, j = J[e] - 1;
          if (i < 0 || j < 0 || i >= n || j >= n) {
            code = -13;
          } else if (k == 0) {
            tother[t].push_back(e);
          } else {
            const long b = (long)q * m + (k - 1);
            ++cnt[b];
            dg[b] += (i == j);
          }
        } else {
          const int size = blockStruct[l] < 0 ? -blockStruct[l] : 1;
          if (I[e] != J[e]) code = -14;
          else if (I[e] < 1 || I[e] > size) code = -13;
          else tother[t].push_back(e);
        }
        if (code) {
          terr[t] = e;
          tcode[t] = code;
          return;
        }
      }
    });
  }
  // the chunks are in input order: the first chunk with an error holds the
  // first erroneous record
  for (int t = 0; t < nt; ++t) {
    if (terr[t] < 0) continue;
    if (tcode[t] == -12) return fail(c, -12, "record out of range");
    if (tcode[t] == -14) return fail(c, -14, "LP entry must be diagonal");
    return fail(c, -13, "index out of range");
  }
  for (int t = 0; t < nt; ++t)
    for (long e : tother[t]) {
      const int k = K[e], l = L[e] - 1;
      if (c->blockType[l] == 1) {
        const int q = c->blockNumber[l], n = c->sdpN[q];
        const int i = I[e] - 1, j = J[e] - 1;
        double* Cq = hC.data() + c->sdpOff[q];
        Cq[i + (size_t)n * j] += -V[e];
        if (i != j) Cq[j + (size_t)n * i] += -V[e];
      } else {
        const int t2 = c->blockNumber[l] + I[e] - 1;
        if (k == 0) {
          hC[c->lpOff + t2] += -V[e];
        } else {
          auto& le = lpent[k - 1];
          bool found = false;
          for (auto& pr : le)
            if (pr.first == t2) { pr.second += V[e]; found = true; }
          if (!found) le.push_back({t2, V[e]});
        }
      }
    }
  tother.clear();
  // bucket sizes, diagonal counts and starts; each chunk's write offsets
  std::vector<long> bstart(nb + 1, 0);
  std::vector<int> bcnt(nb), bdg(nb);
  for (long b = 0; b < nb; ++b) {
    long sz = 0, d = 0;
    for (int t = 0; t < nt; ++t) { sz += tcnt[t][b]; d += tdg[t][b]; }
    if (sz > INT_MAX) return fail(c, -15, "more than 2^31 - 1 SDP entries");
    bcnt[b] = (int)sz;
    bdg[b] = (int)d;
    bstart[b + 1] = bstart[b] + sz;
  }
  const long nentAll = bstart[nb];
  if (nentAll > INT_MAX) return fail(c, -15, "more than 2^31 - 1 SDP entries");
  for (long b = 0; b < nb; ++b) {
    long o = bstart[b];
    for (int t = 0; t < nt; ++t) {
      const int x = tcnt[t][b];
      tcnt[t][b] = (int)o;
      o += x;
    }
  }
  tdg.clear();
  std::unique_ptr<int[]> ber(new int[nentAll ? nentAll : 1]), bec(new int[nentAll ? nentAll : 1]);
  std::unique_ptr<double[]> bev(new double[nentAll ? nentAll : 1]);
  int* h_er = ber.get();
  int* h_ec = bec.get();
  double* h_ev = bev.get();
  {
    const int* bType = c->blockType.data();
    const int* bNum = c->blockNumber.data();
    par_chunks(nt, nnz, [&](int t, long b0, long b1) {
      int* off = tcnt[t].data();
      for (long e = b0; e < b1; ++e) {
        const int k = K[e];
        if (k == 0) continue;
        const int l = L[e] - 1;
        if (bType[l] != 1) continue;
        const int p = off[(long)bNum[l] * m + (k - 1)]++;
        h_er[p] = I[e] - 1;
        h_ec[p] = J[e] - 1;
        h_ev[p] = V[e];
      }
    });
  }
  tcnt.clear();

  // per-block tables: changeToDense, index, formula (sdpa_newton.cpp:515-695)
  c->blk.assign(nS, BlockTable());
  c->nF[0] = c->nF[1] = c->nF[2] = 0;
  std::vector<int> maxNCon(1, 0);
  Arena arena;  // the blocks' small tables, uploaded after the loop
  for (int q = 0; q < nS; ++q) {
    BlockTable& T = c->blk[q];
    T.n = c->sdpN[q];
    T.off = c->sdpOff[q];
    const int n = T.n;
    std::vector<int> con, ecnt, nzeff, isdense;
    for (int k = 0; k < m; ++k) {
      const long bk = (long)q * m + k;
      if (bcnt[bk] == 0) continue;
      con.push_back(k);
      const int cnt = bcnt[bk];
      const bool dense = (double)cnt >= (double)n * n * 0.20;  // sdpa_struct.cpp:518
      const int ne = 2 * cnt - bdg[bk];  // effective entries: off-diagonal ones count twice
      ecnt.push_back(cnt);
      nzeff.push_back(dense ? n * n : ne);
      isdense.push_back(dense ? 1 : 0);
    }
    T.ncon = (int)con.size();
    const int nc = T.ncon;
    // upNonZeroCount (lines 571-596)
    std::vector<double> up(nc);
    {
      // sort-based count: up_j = jnz + sum_{inz < jnz} inz + sum_{inz == jnz, i > j} inz
      std::vector<int> order(nc);
      std::iota(order.begin(), order.end(), 0);
      std::sort(order.begin(), order.end(), [&](int a, int b) {
        return nzeff[a] != nzeff[b] ? nzeff[a] < nzeff[b] : a < b;
      });
      // prefix sums over groups of equal nzeff
      double below = 0.0;
      size_t g0 = 0;
      while (g0 < order.size()) {
        size_t g1 = g0;
        while (g1 < order.size() && nzeff[order[g1]] == nzeff[order[g0]]) ++g1;
        const double z = nzeff[order[g0]];
        // within the group (sorted by index), members after position p have larger index
        for (size_t p = g0; p < g1; ++p) {
          const double larger = (double)(g1 - p - 1);
          up[order[p]] = z + below + larger * z;
        }
        below += z * (double)(g1 - g0);
        g0 = g1;
      }
    }
    std::vector<int> form(nc), dpos(nc, -1), heavy, f1own, dcon;
    std::vector<int> estart(nc);
    for (int k = 0; k < nc; ++k) {
      const double jnz = nzeff[k], nn = n, u = up[k];
      const double f1 = kappa * nn * jnz + nn * nn * nn + kappa * u;
      const double f2 = kappa * nn * jnz + kappa * (nn + 1) * u;
      const double f3 = kappa * (2 * kappa * jnz + 1) * u;
      int f;
      if (isdense[k]) f = (f1 < f2) ? 1 : 2;
      else if (f1 < f2 && f1 < f3) f = (nn <= 200 && 2.0 * nn >= jnz) ? 3 : 1;
      else if (f2 < f3) f = 2;
      else f = 3;
      c->nF[f - 1]++;  // the reference's choice, reported by sdpa_gpu_problem_info
      // The formula a pair is evaluated by: the reference's, except that a
      // sparse F2 owner with at most kLightHost entries in a block of
      // n <= kMultiN takes the F3 gather.  calF2 and calF3 evaluate the same
      // sum A_i . (X A_j Z^-1) in other orders (sdpa_newton.cpp:905-936,
      // 1319-1403); F2 here would run through G_j = X A_j Z^-1 as two n^3
      // products per owner, a batched GEMM launch per block.
      if (f == 2 && !isdense[k] && ecnt[k] <= kLightHost && n <= kMultiN) f = 3;
      form[k] = f;
      if (f == 3) T.nF3++;
      else { T.nF1++; f1own.push_back(k); }
      estart[k] = (int)bstart[(long)q * m + con[k]];
      if (isdense[k]) { dpos[k] = T.ndense++; dcon.push_back(con[k]); }
      if (!isdense[k] && ecnt[k] > kLightHost) heavy.push_back(k);
    }
    T.nheavy = (int)heavy.size();
    {
      // the column-block F3 kernel stages the two rows of X and Z^-1 a
      // one-entry column reads; it pays when nearly every constraint of the
      // block has one entry (constraints spread over many small blocks with
      // a few entries each take the pair kernel)
      long one = 0;
      for (int k = 0; k < nc; ++k) one += ecnt[k] == 1;
      T.f3cols = (n <= 256 && nc >= 64 && one * 10 >= (long)nc * 9) ? 1 : 0;
    }
    arena.add(&T.d_con, con);
    arena.add(&T.d_estart, estart);
    arena.add(&T.d_ecnt, ecnt);
    arena.add(&T.d_nzeff, nzeff);
    arena.add(&T.d_form, form);
    arena.add(&T.d_dpos, dpos);
    arena.add(&T.d_heavy, heavy);
    arena.add(&T.d_f1own, f1own);
    arena.add(&T.d_dcon, dcon);
    if (n <= 256) {  // entry positions fit the record's 8-bit fields
      std::vector<int> rm(4 * (size_t)nc);
      std::vector<double> rv(nc, 0.0);
      for (int k = 0; k < nc; ++k) {
        const int st = estart[k];
        const bool one = ecnt[k] == 1;
        rm[4 * k] = con[k];
        rm[4 * k + 1] = st;
        rm[4 * k + 2] = nzeff[k];
        rm[4 * k + 3] = std::min(ecnt[k], kLightHost + 1) | (form[k] << 4) |
                        (one ? (h_er[st] << 8) | (h_ec[st] << 16) : 0);
        if (one) rv[k] = h_ev[st];
      }
      arena.add(&T.d_rmeta, rm);
      arena.add(&T.d_rval, rv);
    }
    T.h_con = con;
    T.h_form = form;
    T.h_dpos = dpos;
    T.h_f1own = f1own;
    T.h_ecnt = ecnt;
    T.h_estart = estart;
    T.h_nzeff = nzeff;
    if (T.ndense > 0) {
      std::vector<int> dl(T.ndense, 0);
      for (int k = 0; k < nc; ++k)
        if (dpos[k] >= 0) dl[dpos[k]] = k;
      arena.add(&T.d_dlist, dl);
      // one column of the pool per dense A_k (SparseMatrix::changeToDense,
      // sdpa_struct.cpp:510-555: a later duplicate entry overwrites an
      // earlier one), the columns filled by host threads
      const size_t nn2 = (size_t)n * n;
      if (T.d_dense) SDPA_HIP_CHECK(hipFree(T.d_dense));
      T.d_dense = nullptr;
      SDPA_HIP_CHECK(hipMalloc((void**)&T.d_dense, sizeof(double) * nn2 * T.ndense));
      // columns filled into the pinned staging ring chunk by chunk, each
      // chunk's copy running while the next one is filled
      if (int e = ensure_stage(c)) return e;
      const long per = std::max<long>(1L, (long)(kStageBytes / (sizeof(double) * nn2)));
      if (nn2 * sizeof(double) > kStageBytes) {
        // a column larger than a staging buffer: one pageable column at a time
        std::vector<double> A(nn2);
        for (int p = 0; p < T.ndense; ++p) {
          fill_dense_column(A.data(), dl[p], n, estart, ecnt, h_er, h_ec, h_ev);
          SDPA_HIP_CHECK(hipMemcpy(T.d_dense + nn2 * p, A.data(), sizeof(double) * nn2,
                                   hipMemcpyHostToDevice));
        }
      } else {
        int buf = 0;
        for (long p0 = 0; p0 < T.ndense; p0 += per, buf = (buf + 1) % kStageBufs) {
          const long p1 = std::min<long>(T.ndense, p0 + per);
          SDPA_HIP_CHECK(hipEventSynchronize(c->ev_stage[buf]));  // its previous copy is done
          double* st = static_cast<double*>(c->h_stage[buf]);
          const int ntp = (int)std::min<long>(p1 - p0, loader_threads((long)(nn2 * (p1 - p0))));
          par_chunks(ntp, p1 - p0, [&](int, long q0, long q1) {
            for (long q = q0; q < q1; ++q)
              fill_dense_column(st + nn2 * q, dl[p0 + q], n, estart, ecnt, h_er, h_ec, h_ev);
          });
          SDPA_HIP_CHECK(hipMemcpyAsync(T.d_dense + nn2 * p0, st, sizeof(double) * nn2 * (p1 - p0),
                                        hipMemcpyHostToDevice, c->stream));
          SDPA_HIP_CHECK(hipEventRecord(c->ev_stage[buf], c->stream));
        }
        SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
      }
    }
  }
  if (int e = arena.commit(&c->d_blkarena)) return e;
  // The device reads the entry pool only for sparse (not changeToDense)
  // constraints: the F3 kernels, the sparse partners of F1 owners and the
  // densify kernel.  Where every constraint of every block is dense (the
  // dense pool holds them all) it stays on the host.
  bool needEntries = false;
  for (const BlockTable& T : c->blk) needEntries = needEntries || T.ndense < T.ncon;
  const size_t nup = needEntries ? (size_t)nentAll : 0;
  if (int e = upload_raw(c, &c->d_er, h_er, nup)) return e;
  if (int e = upload_raw(c, &c->d_ec, h_ec, nup)) return e;
  if (int e = upload_raw(c, &c->d_ev, h_ev, nup)) return e;
  c->nent = (long)nup;

  // constraint-major sparse terms (dense blocks excluded) and position-major transpose
  std::vector<int> cptr(m + 1, 0);
  std::vector<long> cp1, cp2;
  std::vector<double> cv;
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < nS; ++q) {
      const long b = (long)q * m + k;
      if (bcnt[b] == 0) continue;
      const int n = c->sdpN[q];
      if ((double)bcnt[b] >= (double)n * n * 0.20) continue;  // dense: handled by the pool
      for (long e = bstart[b]; e < bstart[b] + bcnt[b]; ++e) {
        cp1.push_back(c->sdpOff[q] + h_er[e] + (long)n * h_ec[e]);
        cp2.push_back(c->sdpOff[q] + h_ec[e] + (long)n * h_er[e]);
        cv.push_back(h_ev[e]);
      }
    }
    for (auto& pr : lpent[k]) {
      cp1.push_back(c->lpOff + pr.first);
      cp2.push_back(c->lpOff + pr.first);
      cv.push_back(pr.second);
    }
    cptr[k + 1] = (int)cp1.size();
  }
  {
    struct PT { long p; int con; double v; };
    std::vector<PT> pts;
    pts.reserve(cp1.size() * 2);
    for (int k = 0; k < m; ++k)
      for (int e = cptr[k]; e < cptr[k + 1]; ++e) {
        pts.push_back(PT{cp1[e], k, cv[e]});
        if (cp2[e] != cp1[e]) pts.push_back(PT{cp2[e], k, cv[e]});
      }
    // grouped by position, in (constraint, entry) order within a position
    std::vector<long> ppos, pptr(1, 0);
    std::vector<int> pcon;
    std::vector<double> pval;
    const size_t npt = pts.size();
    if ((size_t)c->total <= 8 * npt + ((size_t)1 << 20) && npt < ((size_t)1 << 31)) {
      // a counting sort over the packed positions (stable)
      std::vector<int> start((size_t)c->total + 1, 0);
      for (const PT& x : pts) ++start[x.p + 1];
      for (long q = 0; q < c->total; ++q) start[q + 1] += start[q];
      std::vector<int> at(start.begin(), start.end() - 1);
      pcon.resize(npt);
      pval.resize(npt);
      for (const PT& x : pts) {
        const int w = at[x.p]++;
        pcon[w] = x.con;
        pval[w] = x.v;
      }
      for (long q = 0; q < c->total; ++q)
        if (start[q + 1] > start[q]) {
          ppos.push_back(q);
          pptr.push_back(start[q + 1]);
        }
    } else {
      std::stable_sort(pts.begin(), pts.end(), [](const PT& a, const PT& b) { return a.p < b.p; });
      for (size_t t = 0; t < npt; ++t) {
        if (t == 0 || pts[t].p != pts[t - 1].p) {
          if (t > 0) pptr.push_back((long)t);
          ppos.push_back(pts[t].p);
        }
        pcon.push_back(pts[t].con);
        pval.push_back(pts[t].v);
      }
      if (npt) pptr.push_back((long)npt);
    }
    c->npos = (long)ppos.size();
    if (int e = upload(c, &c->d_ppos, ppos)) return e;
    if (int e = upload(c, &c->d_pptr, pptr)) return e;
    if (int e = upload(c, &c->d_pcon, pcon)) return e;
    if (int e = upload(c, &c->d_pval, pval)) return e;
  }
  if (int e = upload(c, &c->d_cptr, cptr)) return e;
  if (int e = upload(c, &c->d_cp1, cp1)) return e;
  if (int e = upload(c, &c->d_cp2, cp2)) return e;
  if (int e = upload(c, &c->d_cv, cv)) return e;

  // LP coordinates (InputData::initialize_index_LP)
  {
    std::vector<int> lpptr(c->nLP + 1, 0), lpcon;
    std::vector<double> lpval;
    std::vector<std::vector<std::pair<int, double>>> byc(c->nLP);
    for (int k = 0; k < m; ++k)
      for (auto& pr : lpent[k]) byc[pr.first].push_back({k, pr.second});
    for (int t = 0; t < c->nLP; ++t) {
      for (auto& pr : byc[t]) { lpcon.push_back(pr.first); lpval.push_back(pr.second); }
      lpptr[t + 1] = (int)lpcon.size();
    }
    if (int e = upload(c, &c->d_lpptr, lpptr)) return e;
    if (int e = upload(c, &c->d_lpcon, lpcon)) return e;
    if (int e = upload(c, &c->d_lpval, lpval)) return e;
    // Chordal::ordering_bMat dense decision (sdpa_chordal.cpp:376-410)
    c->schurDense = 0;
    if (m <= 100) c->schurDense = 1;
    const double thr = m * std::sqrt(0.70);
    for (int q = 0; q < nS; ++q)
      if (c->blk[q].ncon > thr) c->schurDense = 1;
    for (int t = 0; t < c->nLP; ++t)
      if ((double)byc[t].size() > thr) c->schurDense = 1;
    if (!c->schurDense) {
      // aggregate pattern of B (sdpa_chordal.cpp:404-410): pairs of
      // constraints sharing an SDP block or an LP coordinate, plus the
      // diagonal; dense when the full count exceeds 0.7 m^2
      const size_t words = ((size_t)m + 63) / 64;
      std::vector<uint64_t> pat((size_t)m * words, 0ull), gs(words);
      // row i of the pattern gains every member j <= i of the group g: the
      // group's bit set, whole words below i's word and i's word masked
      auto mark = [&](const std::vector<int>& g) {
        std::fill(gs.begin(), gs.end(), 0ull);
        for (int i : g) gs[i >> 6] |= 1ull << (i & 63);
        for (int i : g) {
          uint64_t* row = pat.data() + (size_t)i * words;
          const int wi = i >> 6, bi = i & 63;
          for (int w = 0; w < wi; ++w) row[w] |= gs[w];
          row[wi] |= gs[wi] & (bi == 63 ? ~0ull : (2ull << bi) - 1);
        }
      };
      for (int q = 0; q < nS; ++q) mark(c->blk[q].h_con);
      std::vector<int> g;
      for (int t = 0; t < c->nLP; ++t) {
        g.clear();
        for (auto& pr : byc[t]) g.push_back(pr.first);
        mark(g);
      }
      long lower = 0;
      for (int i = 0; i < m; ++i) {
        pat[(size_t)i * words + (i >> 6)] |= 1ull << (i & 63);
        for (size_t w = 0; w < words; ++w) lower += __builtin_popcountll(pat[(size_t)i * words + w]);
      }
      if ((double)(2 * lower - m) > 0.70 * (double)m * (double)m) c->schurDense = 1;
    }
    if (!c->schurDense) {
      // the sparse branch: ordering and tile analysis (sparse_schur.cpp);
      // the tile factorization is taken when its flops (x 1.15) stay below
      // 0.85 of the dense m^3 / 3, the reference's choice between its sparse
      // solver and the dense one (sdpa_chordal.cpp:497-514); otherwise the
      // dense factorization runs on the same shifted matrix
      std::vector<std::vector<int>> cliques;
      for (int q = 0; q < nS; ++q) cliques.push_back(c->blk[q].h_con);
      for (int t = 0; t < c->nLP; ++t) {
        std::vector<int> g;
        for (auto& pr : byc[t]) g.push_back(pr.first);
        cliques.push_back(std::move(g));
      }
      const int ar = sparse_schur_analyse(m, cliques, 4.0e9, &c->ts);
      const double dense_cost = (double)m * m * m / 3.0;
      if (ar == 0 && c->ts.flops * 1.15 <= 0.85 * dense_cost) {
        if (int e = sparse_schur_upload(&c->ts)) return e;
        c->schurSparse = 1;
      } else {
        sparse_schur_free(&c->ts);
      }
    }
  }
  c->bmatStoreFirst = (nS == 1 && c->nLP == 0 && c->blk[0].ncon == m && c->blk[0].nF1 == 0);
  c->f3multi = 0;
  {
    // the blocks it takes: n <= kMultiN (the others keep their per-block
    // kernels, which then add to B)
    long nf3 = 0;
    for (int q = 0; q < nS; ++q)
      if (c->blk[q].n <= kMultiN) nf3 += c->blk[q].nF3;
    if (nS > 1 && nS < (1 << 20) && nf3 > 0 && m <= kF3MultiMaxM) {
      // per block its entry records in chunks of 64, no constraint's entries
      // straddling two chunks; constraints with more than kLightHost entries
      // (heavy) are not listed
      std::vector<int> mrec;  // int4 each
      std::vector<long> rec0(nS, 0);
      std::vector<int> nchunk(nS, 0);
      std::vector<std::vector<int>> first(nS);
      int pmax = 0, nmax = 0;
      auto pad_to = [&](long n4) {
        while ((long)mrec.size() < 4 * n4) mrec.push_back(0);
      };
      for (int q = 0; q < nS; ++q) {
        const BlockTable& T = c->blk[q];
        if (T.n > kMultiN) continue;
        pmax = std::max(pmax, T.n * (T.n + 1) / 2);
        nmax = std::max(nmax, T.n);
        rec0[q] = (long)mrec.size() / 4;
        long fill = 0;
        first[q].assign(T.ncon, -1);
        for (int pos = 0; pos < T.ncon; ++pos) {
          const int ci = T.h_ecnt[pos];
          if (ci < 1 || ci > kLightHost) continue;
          if ((fill & 63) + ci > 64) fill = (fill | 63) + 1;
          pad_to(rec0[q] + fill);
          first[q][pos] = (int)fill;
          const int st = T.h_estart[pos];
          for (int e = 0; e < ci; ++e) {
            const int lo = std::min(h_er[st + e], h_ec[st + e]);
            const int hi = std::max(h_er[st + e], h_ec[st + e]);
            const int meta = T.h_con[pos] | ((hi * (hi + 1) / 2 + lo) << 13) |
                             ((ci - 1 - e) << 25) | ((e == 0 ? 1 : 0) << 28) |
                             ((T.h_form[pos] == 3 ? 1 : 0) << 29);
            int vw[2];
            std::memcpy(vw, &h_ev[st + e], sizeof(vw));
            mrec.insert(mrec.end(), {meta, T.h_nzeff[pos], vw[0], vw[1]});
          }
          fill += ci;
        }
        fill = (fill + 63) & ~63L;
        pad_to(rec0[q] + fill);
        nchunk[q] = (int)(fill / 64);
      }
      std::vector<int> cnt(m + 1, 0);
      for (int q = 0; q < nS; ++q)
        for (int pos = 0; pos < (int)first[q].size(); ++pos)
          if (first[q][pos] >= 0) ++cnt[c->blk[q].h_con[pos] + 1];
      for (int k = 0; k < m; ++k) cnt[k + 1] += cnt[k];
      constexpr int kW = 4 * kVisInt4;  // ints per visit record
      std::vector<int> jent(kW * (size_t)cnt[m], 0), fill(cnt.begin(), cnt.end() - 1);
      for (int q = 0; q < nS; ++q) {
        const BlockTable& T = c->blk[q];
        for (int pos = 0; pos < (int)first[q].size(); ++pos) {
          if (first[q][pos] < 0) continue;
          int* r = &jent[kW * (size_t)fill[T.h_con[pos]]++];
          const int st = T.h_estart[pos];
          r[0] = (int)rec0[q];
          r[1] = nchunk[q];
          r[2] = T.n | (std::min(T.h_ecnt[pos], 15) << 8) | ((T.h_form[pos] == 3 ? 1 : 0) << 16);
          r[3] = first[q][pos] >> 6;
          r[4] = T.h_nzeff[pos];
          r[5] = st;
          std::memcpy(&r[6], &T.off, sizeof(long));
          for (int e = 0; e < std::min(T.h_ecnt[pos], 4); ++e) {
            r[8 + e / 2] |= (int)((unsigned)(h_er[st + e] | (h_ec[st + e] << 8)) << (16 * (e & 1)));
            std::memcpy(&r[12 + 2 * e], &h_ev[st + e], sizeof(double));
          }
        }
      }
      if ((long)mrec.size() / 4 > (long)INT_MAX) return -1;
      if (int e = upload(c, &c->d_mrec, mrec)) return e;
      if (int e = upload(c, &c->d_jptr, cnt)) return e;
      if (int e = upload(c, &c->d_jent, jent)) return e;
      if (int e = f3multi_prepare(m, pmax, nmax)) return e;
      c->f3m_pmax = pmax;
      c->f3m_nmax = nmax;
      c->f3multi = 1;
    }
  }

  std::vector<double> hb(cvec, cvec + m);
  if (int e = upload(c, &c->d_b, hb)) return e;
  if (int e = upload(c, &c->d_C, hC)) return e;
  if (int e = upload(c, &c->d_sdpN, c->sdpN)) return e;
  if (int e = upload(c, &c->d_sdpOff, c->sdpOff)) return e;
  const size_t T = c->total;
  double** packed[] = {&c->d_X, &c->d_Z, &c->d_invCholX, &c->d_invCholZ, &c->d_Zinv, &c->d_LX,
                       &c->d_LZ, &c->d_Dx, &c->d_Dz, &c->d_rzinv, &c->d_xrdzinv, &c->d_w2,
                       &c->d_dmat, &c->d_X0, &c->d_Z0};
  for (double** p : packed)
    if (int e = dalloc(p, T)) return e;
  if (int e = dalloc(&c->d_w1, 2 * T)) return e;
  if (int e = dalloc(&c->d_w3, 2 * T)) return e;
  double** mvec[] = {&c->d_y, &c->d_dy, &c->d_g, &c->d_pvec};
  for (double** p : mvec)
    if (int e = dalloc(p, m)) return e;
  if (int e = dalloc(&c->d_B, (size_t)m * m)) return e;
  if (int e = dalloc(&c->d_Binv, (size_t)((m + 63) / 64) * 4096)) return e;
  if (int e = dalloc(&c->d_Winv, (size_t)((m + 255) / 256) * 65536)) return e;
  if (int e = dalloc(&c->d_red, (size_t)kMaxRed * 4096 + 64)) return e;
  if (int e = dalloc(&c->d_eig, 2 * (size_t)(nS + 1))) return e;
  if (c->h_eig) (void)hipHostFree(c->h_eig);
  c->h_eig = nullptr;
  SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_eig, sizeof(double) * 2 * (nS + 1),
                               hipHostMallocDefault));
  SDPA_HIP_CHECK(hipMalloc(&c->d_info, sizeof(int) * (2 * nS + 2)));
  if (2 * nS + 2 > 64) {
    (void)hipHostFree(c->h_info);
    SDPA_HIP_CHECK(hipHostMalloc((void**)&c->h_info, sizeof(int) * (2 * nS + 2), hipHostMallocDefault));
  }
  c->loaded = true;
  return 0;
}

int sdpa_gpu_problem_info(sdpa_gpu_ctx* h, long* out, int cap) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  long v[16] = {c->m, c->nBlock, c->nSDP, c->nLP, c->nDim, c->total, c->nF[0], c->nF[1],
                c->nF[2], c->schurDense, c->bmatStoreFirst ? 1 : 0, c->nSDP, c->schurSparse,
                c->schurSparse ? c->ts.ntiles : 0, c->schurSparse ? c->ts.T : 0,
                c->schurSparse ? c->ts.nlev : 0};
  int n = cap < 16 ? cap : 16;
  for (int t = 0; t < n; ++t) out[t] = v[t];
  return n;
}

int sdpa_gpu_block_size(sdpa_gpu_ctx* h, int q) {
  if (!h || !h->c.loaded || q < 0 || q >= h->c.nSDP) return -1;
  return h->c.sdpN[q];
}

static double* buffer_ptr(Ctx* c, int which, long* len) {
  switch (which) {
    case 0: *len = c->total; return c->d_X;
    case 1: *len = c->total; return c->d_Z;
    case 2: *len = c->m; return c->d_y;
    case 3: *len = c->total; return c->d_Zinv;
    case 4: *len = c->total; return c->d_invCholX;
    case 5: *len = c->total; return c->d_invCholZ;
    case 6: *len = c->total; return c->d_Dx;
    case 7: *len = c->total; return c->d_Dz;
    case 8: *len = c->m; return c->d_dy;
    case 9: *len = c->m; return c->d_g;
    case 10: *len = (long)c->m * c->m; return c->d_B;
    case 11: *len = c->m; return c->d_pvec;
    case 12: *len = c->total; return c->d_dmat;
    default: *len = 0; return nullptr;
  }
}

int sdpa_gpu_get_buffer(sdpa_gpu_ctx* h, int which, double* dst, long len) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  long L;
  double* p = buffer_ptr(c, which, &L);
  if (!p || len != L) return fail(c, -20, "buffer id or length mismatch");
  SDPA_HIP_CHECK(hipMemcpyAsync(dst, p, sizeof(double) * L, hipMemcpyDeviceToHost, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int sdpa_gpu_set_buffer(sdpa_gpu_ctx* h, int which, const double* src, long len) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  long L;
  double* p = buffer_ptr(c, which, &L);
  if (!p || len != L) return fail(c, -20, "buffer id or length mismatch");
  SDPA_HIP_CHECK(hipMemcpyAsync(p, src, sizeof(double) * L, hipMemcpyHostToDevice, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (which == 10) c->bshifted = false;  // a new Schur matrix
  return 0;
}

int sdpa_gpu_set_iterate(sdpa_gpu_ctx* h, const double* X, const double* y, const double* Z) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  if (X) SDPA_HIP_CHECK(hipMemcpyAsync(c->d_X, X, sizeof(double) * c->total, hipMemcpyHostToDevice, c->stream));
  if (y) SDPA_HIP_CHECK(hipMemcpyAsync(c->d_y, y, sizeof(double) * c->m, hipMemcpyHostToDevice, c->stream));
  if (Z) SDPA_HIP_CHECK(hipMemcpyAsync(c->d_Z, Z, sizeof(double) * c->total, hipMemcpyHostToDevice, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int sdpa_gpu_get_iterate(sdpa_gpu_ctx* h, double* X, double* y, double* Z) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  if (X) SDPA_HIP_CHECK(hipMemcpyAsync(X, c->d_X, sizeof(double) * c->total, hipMemcpyDeviceToHost, c->stream));
  if (y) SDPA_HIP_CHECK(hipMemcpyAsync(y, c->d_y, sizeof(double) * c->m, hipMemcpyDeviceToHost, c->stream));
  if (Z) SDPA_HIP_CHECK(hipMemcpyAsync(Z, c->d_Z, sizeof(double) * c->total, hipMemcpyDeviceToHost, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int sdpa_gpu_identity_start(sdpa_gpu_ctx* h, double lambda) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<double> X(c->total, 0.0), IC(c->total, 0.0), ZI(c->total, 0.0);
  for (int q = 0; q < c->nSDP; ++q)
    for (int k = 0; k < c->sdpN[q]; ++k) {
      long d = c->sdpOff[q] + (long)k * c->sdpN[q] + k;
      X[d] = lambda;
      IC[d] = 1.0 / std::sqrt(lambda);
      ZI[d] = 1.0 / lambda;
    }
  for (int t = 0; t < c->nLP; ++t) {
    long d = c->lpOff + t;
    X[d] = lambda;
    IC[d] = 1.0 / std::sqrt(lambda);
    ZI[d] = 1.0 / lambda;
  }
  const size_t B = sizeof(double) * c->total;
  SDPA_HIP_CHECK(hipMemcpy(c->d_X, X.data(), B, hipMemcpyHostToDevice));
  SDPA_HIP_CHECK(hipMemcpy(c->d_Z, X.data(), B, hipMemcpyHostToDevice));
  SDPA_HIP_CHECK(hipMemcpy(c->d_invCholX, IC.data(), B, hipMemcpyHostToDevice));
  SDPA_HIP_CHECK(hipMemcpy(c->d_invCholZ, IC.data(), B, hipMemcpyHostToDevice));
  SDPA_HIP_CHECK(hipMemcpy(c->d_Zinv, ZI.data(), B, hipMemcpyHostToDevice));
  SDPA_HIP_CHECK(hipMemset(c->d_y, 0, sizeof(double) * c->m));
  return 0;
}

int sdpa_gpu_keep_initial_point(sdpa_gpu_ctx* h) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  const size_t B = sizeof(double) * c->total;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->d_X0, c->d_X, B, hipMemcpyDeviceToDevice, c->stream));
  SDPA_HIP_CHECK(hipMemcpyAsync(c->d_Z0, c->d_Z, B, hipMemcpyDeviceToDevice, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// -- timing helpers ----------------------------------------------------------
static void tic(Ctx* c) {
  if (c->timing) (void)hipEventRecord(c->ev[0], c->stream);
}
static double toc(Ctx* c) {
  if (!c->timing) return 0.0;
  (void)hipEventRecord(c->ev[1], c->stream);
  (void)hipEventSynchronize(c->ev[1]);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
  return ms;
}

int sdpa_gpu_inv_chol(sdpa_gpu_ctx* h) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  int r = inverse_chol(c);
  c->t_inv += toc(c);
  return r;
}

int sdpa_gpu_make_bmat(sdpa_gpu_ctx* h, int precision) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  if (precision != 0 && precision != 1) return fail(c, -30, "precision must be 0 or 1");
  tic(c);
  int r = make_bmat(c, precision);
  c->bshifted = false;
  c->t_bmat += toc(c);
  c->n_bmat++;
  return r;
}

int sdpa_gpu_factor_bmat(sdpa_gpu_ctx* h, int precision) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  if (precision != 0 && precision != 1) return fail(c, -30, "precision must be 0 (FP64) or 1 (mixed)");
  tic(c);
  c->factor_mixed = 0;
  if (c->schurSparse) {
    // Chordal::factorizeSchur (sdpa_chordal.cpp:526-571): bMat + 1e-10 I
    // factored by the sparse branch (FP64 at either precision); the shift is
    // added as the tiles are gathered, d_B keeps the assembled matrix
    const int r = sparse_schur_factor(c->stream, &c->ts, c->d_B, 1.0e-10);
    c->t_chol += toc(c);
    c->n_chol++;
    return r;
  }
  if (!c->schurDense && !c->bshifted) {
    // Chordal::factorizeSchur (sdpa_chordal.cpp:526-534): where the reference
    // leaves the dense path it factors bMat + 1e-10 I (PLUS_ADJUST_DIAGONAL)
    // with its sparse solver; the same shifted matrix is factored here, once
    // per assembled matrix (a refactorization after a stalled mixed solve
    // sees it already shifted), and the mixed path refines against it
    hipLaunchKernelGGL(k_shift_diag, dim3((c->m + 255) / 256), dim3(256), 0, c->stream, c->d_B,
                       c->m, 1.0e-10);
    c->bshifted = true;
  }
  if (precision == 1) {
    const int r = factor_mixed(c);
    if (r < 0) return r;
    if (r == 0) {
      c->t_chol += toc(c);
      c->n_chol++;
      return 0;
    }
    ++c->mx_fallbacks;  // FP32 pivot failure or size out of range: FP64 below (B intact)
  }
  MatDesc d{c->d_B, c->m, c->m};
  int r = potrf_batched(c->stream, &c->dw, &d, 1, 1, c->d_info, c->d_Binv);
  if (r) return r;
  SDPA_HIP_CHECK(hipMemcpyAsync(c->h_info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (int e = c->dw.collect_trailing()) return e;
  if (c->h_info[0] == 0) {
    r = diag_block_inverses(c->stream, &c->dw, c->d_B, c->m, c->m, c->d_Winv, c->d_Binv);
    if (r) return r;
  }
  c->t_chol += toc(c);
  c->n_chol++;
  return c->h_info[0];
}

int sdpa_gpu_solve_bmat(sdpa_gpu_ctx* h, int refine) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  (void)refine;
  tic(c);
  if (c->factor_mixed) {
    const int r = solve_mixed(c);
    c->t_solve += toc(c);
    c->n_solve++;
    return r;
  }
  if (c->schurSparse) {
    // Chordal::solveSchur (sdpa_chordal.cpp:573-579)
    const int r = sparse_schur_solve(c->stream, &c->ts, c->d_g, c->d_dy);
    c->t_solve += toc(c);
    c->n_solve++;
    return r;
  }
  int r = chol_solve(c->stream, &c->dw, c->d_B, c->m, c->m, c->d_Winv, c->d_dy, c->d_g);
  c->t_solve += toc(c);
  c->n_solve++;
  return r;
}

int sdpa_gpu_min_eig(sdpa_gpu_ctx* h, double* lamX, double* lamZ) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  int r = min_eig(c, lamX, lamZ);
  c->t_eig += toc(c);
  return r;
}

int sdpa_gpu_residuals(sdpa_gpu_ctx* h, double* nP, double* nD) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  const int r = residuals(c, nP, nD);
  c->t_res += toc(c);
  return r;
}

int sdpa_gpu_make_gvec(sdpa_gpu_ctx* h, int corrector, double target, int infeasible) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  const int r = make_gvec(c, corrector, target, infeasible);
  c->t_gvec += toc(c);
  return r;
}

int sdpa_gpu_make_direction(sdpa_gpu_ctx* h, int infeasible) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  const int r = make_direction(c, infeasible);
  c->t_dir += toc(c);
  return r;
}

// after a synchronization: the failure flag of the last cooperative Schur
// solve (copied to pinned memory behind it) is current
static int coop_solve_check(sdpa_gpu_ctx* h) {
  if (h->c.dw.h_err && *h->c.dw.h_err) {
    if (int e = h->c.dw.clear_err(h->c.stream)) return e;
    return fail(&h->c, -60, "a cooperative Schur solve gave up waiting (work-groups not co-resident)");
  }
  return 0;
}

int sdpa_gpu_min_eig_dots(sdpa_gpu_ctx* h, double* lamX, double* lamZ, int nq, const int* codes,
                          double* out) {
  if (!h || !h->c.loaded) return -1;
  if (nq < 0 || nq > 32) return -2;
  Ctx* c = &h->c;
  tic(c);
  if (int e = min_eig_enqueue(c, kSlotLpMin)) return e;
  c->t_eig += toc(c);
  tic(c);
  if (nq > 0)
    if (int e = dots_enqueue(c, nq, codes, 0)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->t_dots += toc(c);
  min_eig_finish(c, kSlotLpMin, lamX, lamZ);
  for (int k = 0; k < nq; ++k) out[k] = c->h_red[k];
  return coop_solve_check(h);
}

int sdpa_gpu_update_end(sdpa_gpu_ctx* h, int nq, const int* codes, double* out, double* nP,
                        double* nD) {
  if (!h || !h->c.loaded) return -1;
  if (nq < 0 || nq > 32) return -2;
  Ctx* c = &h->c;
  tic(c);
  if (int e = inverse_chol_enqueue(c)) return e;
  c->t_inv += toc(c);
  tic(c);
  if (nq > 0)
    if (int e = dots_enqueue(c, nq, codes, 0)) return e;
  c->t_dots += toc(c);
  tic(c);
  if (int e = residuals_enqueue(c, kSlotResid)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->t_res += toc(c);
  if (inverse_chol_status(c)) return 1;
  for (int k = 0; k < nq; ++k) out[k] = c->h_red[k];
  *nP = c->h_red[kSlotResid];
  *nD = c->h_red[kSlotResid + 1];
  return coop_solve_check(h);
}

int sdpa_gpu_dots(sdpa_gpu_ctx* h, int nq, const int* codes, double* out) {
  if (!h || !h->c.loaded) return -1;
  if (nq <= 0 || nq > 32) return -2;
  tic(&h->c);
  const int r = compute_dots(&h->c, nq, codes, out);
  h->c.t_dots += toc(&h->c);
  // compute_dots synchronizes the stream
  if (r == 0) return coop_solve_check(h);
  return r;
}

int sdpa_gpu_mixed_stats(sdpa_gpu_ctx* h, long* out6) {
  if (!h || !out6) return -1;
  const Ctx* c = &h->c;
  out6[0] = c->mx_factors;
  out6[1] = c->mx_fallbacks;
  out6[2] = c->mx_solves;
  out6[3] = c->mx_refine;
  out6[4] = c->mx_stalls;
  out6[5] = c->factor_mixed;
  return 0;
}

int sdpa_gpu_update(sdpa_gpu_ctx* h, double aP, double aD) {
  if (!h || !h->c.loaded) return -1;
  Ctx* c = &h->c;
  tic(c);
  const int r = update_point(c, aP, aD);
  c->t_upd += toc(c);
  return r;
}

int sdpa_gpu_timing_detail(sdpa_gpu_ctx* h, double* out, int cap) {
  if (!h || !out) return -1;
  const Ctx* c = &h->c;
  const double v[11] = {c->t_bmat, c->t_f1,  c->t_chol, c->t_solve, c->t_eig, c->t_inv,
                        c->t_gvec, c->t_dir, c->t_res,  c->t_upd,   c->t_dots};
  const int n = cap < 11 ? cap : 11;
  for (int t = 0; t < n; ++t) out[t] = v[t];
  return n;
}

int sdpa_gpu_timing(sdpa_gpu_ctx* h, int enable, double* out) {
  if (!h) return -1;
  Ctx* c = &h->c;
  if (out) {
    out[0] = c->t_bmat; out[1] = c->n_bmat;
    out[2] = c->t_chol; out[3] = c->n_chol;
    out[4] = c->t_solve; out[5] = c->n_solve;
    out[6] = c->t_eig; out[7] = 0;
    out[8] = c->t_inv; out[9] = 0;
  }
  if (enable >= 0) {
    c->timing = enable != 0;
    c->t_bmat = c->t_chol = c->t_solve = c->t_eig = c->t_inv = 0;
    c->t_f1 = c->t_gvec = c->t_dir = c->t_res = c->t_upd = c->t_dots = 0;
    c->n_bmat = c->n_chol = c->n_solve = 0;
    c->dw.time_trailing = c->timing;
    c->dw.trail_ms = c->dw.trail_flops = 0.0;
    c->dw.trail_launches = 0;
    c->dw.panel_ms = c->dw.panel_flops = 0.0;
    c->dw.panel_launches = 0;
  }
  return 0;
}

int sdpa_gpu_kernel_timing(sdpa_gpu_ctx* h, double* out8) {
  if (!h || !out8) return -1;
  Ctx* c = &h->c;
  out8[0] = c->dw.trail_ms;
  out8[1] = (double)c->dw.trail_launches;
  out8[2] = c->dw.trail_flops;
  out8[3] = 0.0;
  out8[4] = c->dw.panel_ms;
  out8[5] = (double)c->dw.panel_launches;
  out8[6] = c->dw.panel_flops;
  out8[7] = 0.0;
  return 0;
}

int sdpa_gpu_synchronize(sdpa_gpu_ctx* h) {
  if (!h) return -1;
  SDPA_HIP_CHECK(hipStreamSynchronize(h->c.stream));
  return 0;
}

}  // extern "C"
