// common.h -- shared definitions of the sdpa_amd HIP library (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace sdpa_amd {

#define SDPA_HIP_CHECK(expr)                                                           \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      std::fprintf(stderr, "sdpa_amd: HIP error %s at %s:%d: %s\n", hipGetErrorString(_e), \
                   __FILE__, __LINE__, #expr);                                         \
      return -1000 - (int)_e;                                                          \
    }                                                                                  \
  } while (0)

constexpr int kWave = 64;

// Descriptor of one column-major matrix inside a batch.
struct MatDesc {
  double* ptr;
  int n;    // rows == cols for the square factorizations
  int ld;
};

// Batched GEMM: C = alpha * op(A) * op(B) + beta * C for every item.
struct GemmItem {
  const double* A;
  const double* B;
  double* C;
  int M, N, K;
  int lda, ldb, ldc;
};

// ---- dense kernels (dense.hip) ----
// transA/transB: 0 = N, 1 = T.  lower_only: compute only tiles/elements with
// row >= col of C (C square).  Items may have different shapes.
int gemm_batched(hipStream_t s, const GemmItem* d_items, int nitems, int maxM, int maxN,
                 int transA, int transB, double alpha, double beta, int lower_only);
// Convenience single GEMM (copies one item to a device scratch slot).
struct DenseWork;  // per-context scratch for descriptor arrays
int gemm(hipStream_t s, DenseWork* w, const double* A, int lda, const double* B, int ldb,
         double* C, int ldc, int M, int N, int K, int transA, int transB, double alpha,
         double beta, int lower_only);

// Blocked lower Cholesky of a batch of matrices.  mode 0: LAPACK dpotrf
// semantics (fail at the first pivot that is not > 0); mode 1: SDPA's adjust
// rule (pivot < -1e-6 fails, pivot < 1e-14 becomes 1e100).  d_info[b]
// receives 0 or the 1-based failing column of item b.
// dinv_out (optional): receives inv(L_kk) of every 64x64 diagonal block,
// slot (item, k0/64), 64*64 doubles each, ceil(maxn/64) slots per item.
int potrf_batched(hipStream_t s, DenseWork* w, const MatDesc* h_items, int nitems, int mode,
                  int* d_info, double* dinv_out = nullptr);
// Mixed-precision factorization: A32 (m x m floats, ld m) <- FP32 lower
// Cholesky factor of S B S, S = diag(B_ii^-1/2) into scale; B (FP64, lower
// triangle, ld) is left intact.  Panels are factored in FP64 in P64 (m x 256)
// with SDPA's pivot rule; D64 receives the FP64 256x256 diagonal blocks of
// the factor (ceil(m/256) blocks, ld 256) and dinv their 64x64 inverses.
// *d_info: failing column or 0; *d_bad: a diagonal of B was not positive;
// *d_bmax: bits of max |B_ij|.
int potrf_mixed(hipStream_t s, DenseWork* w, const double* B, int ld, int m, float* A32,
                double* scale, double* P64, double* D64, double* dinv, int* d_info, int* d_bad,
                unsigned long long* d_bmax);
// Sparse-Schur branch: a fill-reducing ordering of the Schur matrix and the
// factorization of the permuted matrix by 64 x 64 tiles, only the tiles of
// the factor's (filled) pattern stored and computed (sparse_schur.cpp for the
// analysis, dense.hip for the kernels).
struct TileSparse {
  int m = 0, T = 0, ntiles = 0, nlev = 0;
  std::vector<int> perm, iperm;   // position in the ordering -> constraint, inverse
  // tiles of L column by column, the diagonal tile first: tile p is (tI[p], tJ[p])
  std::vector<int> colptr, tI, tJ, tile_of;  // tile_of[I * T + J]: tile id or -1
  // schedule: the tile columns by height in the tile elimination tree (a
  // level's columns are independent); per level the off-diagonal tiles to
  // solve and the update tasks, one per target tile with its source pairs
  // (tile ids of L(I, J), L(K, J)) in column order: target t's pairs are
  // upd_src[2 q], upd_src[2 q + 1] for upd_src_ptr[t] <= q < upd_src_ptr[t + 1]
  std::vector<int> lvl_ptr, lvl_cols, trsm_ptr, trsm_tiles;
  std::vector<int> upd_ptr, upd_target, upd_src_ptr, upd_src;
  // a target with many sources: its chunks write partial sums to scratch
  // slots (upd_slot >= 0, else -1: subtracted at once), then a reduction
  // task per level subtracts slots red_slot0 .. + red_nslot in order
  std::vector<int> upd_slot, red_ptr, red_target, red_slot0, red_nslot;
  int max_slots = 0;
  // forward solve: per level the tile rows I updated by its columns, each
  // with its tiles (x_I -= sum L(I, J) x_J over fwd_src[fwd_src_ptr[r] ..])
  std::vector<int> fwd_ptr, fwd_row, fwd_src_ptr, fwd_src;
  double flops = 0.0, ext_nnz = 0.0;  // tile-level estimates
  double* d_tiles = nullptr;  // ntiles x 4096 doubles (column-major tiles)
  double* d_winv = nullptr;   // T x 4096: inverses of the diagonal tiles of L
  double* d_x = nullptr;      // 64 T: solve workspace
  double* d_u = nullptr;      // 64 ntiles: products L(I, J) x_J of the forward solve
  int* d_info = nullptr;      // T: failing column within each diagonal tile (1-based) or 0
  int *d_perm = nullptr, *d_tI = nullptr, *d_tJ = nullptr, *d_colptr = nullptr;
  int *d_lvl_cols = nullptr, *d_trsm_tiles = nullptr, *d_upd_target = nullptr;
  int *d_upd_src_ptr = nullptr, *d_upd_src = nullptr;
  int *d_fwd_row = nullptr, *d_fwd_src_ptr = nullptr, *d_fwd_src = nullptr;
  int *d_upd_slot = nullptr, *d_red_target = nullptr, *d_red_slot0 = nullptr,
      *d_red_nslot = nullptr;
  double* d_part = nullptr;  // max_slots x 4096: partial update sums
};
// Ordering, symbolic factorization and schedule from the cliques of the
// pattern (constraint lists of the SDP blocks and LP coordinates).  Returns
// 0, 1 when the analysis would exceed max_work word operations, 2 on an
// inconsistent symbolic factorization.
int sparse_schur_analyse(int m, const std::vector<std::vector<int>>& cliques, double max_work,
                         TileSparse* ts);
int sparse_schur_upload(TileSparse* ts);  // device copies of the tables and workspaces
void sparse_schur_free(TileSparse* ts);
// Factor P (B + shift I) P^T (B: dense m x m, lower triangle, ld m) by tiles
// with LAPACK's pivot rule (a pivot that is not > 0 fails).  Returns 0, a
// 1-based failing column of the permuted matrix, or < 0 on a device error.
int sparse_schur_factor(hipStream_t s, TileSparse* ts, const double* B, double shift);
// dy := (P^T L L^T P)^-1 g with the tile factor, level by level.
int sparse_schur_solve(hipStream_t s, TileSparse* ts, const double* g, double* dy);

// W = L^-1 for a batch of lower factors (upper parts of W zeroed).
// dinv_pre (optional): the 64x64 diagonal inverses potrf_batched produced for
// the same batch (slot (item, k0/64)); otherwise they are computed here.
int trtri_batched(hipStream_t s, DenseWork* w, const MatDesc* h_L, const MatDesc* h_W,
                  int nitems, const double* dinv_pre = nullptr);
// Inverses of the 256x256 diagonal blocks of the m x m lower factor L into
// Winv (ceil(m/256) blocks of 256*256 doubles, ld 256).  dinv64 (optional):
// potrf_batched's 64x64 diagonal inverses of L.
int diag_block_inverses(hipStream_t s, DenseWork* w, const double* L, int ld, int m,
                        double* Winv, const double* dinv64 = nullptr);
// x := L^-T L^-1 x (two triangular solves with the lower factor L, m x m),
// using the block inverses from diag_block_inverses; rhs (optional): the
// right-hand side read from there instead of x, left unchanged.
int chol_solve(hipStream_t s, DenseWork* w, const double* L, int ld, int m, const double* Winv,
               double* x, const double* rhs = nullptr);
// The same with a float factor (mixed precision); returns 1 if the
// pipelined kernels do not apply to this size (m <= 512 included).
int chol_solve_f32(hipStream_t s, DenseWork* w, const float* L, int ld, int m, const double* Winv,
                   double* x);
bool chol_solve_f32_supported(int m, int ld);

struct DenseWork {
  double* d_dinv = nullptr;  // inverses of 64x64 diagonal blocks
  size_t cap_dinv = 0;
  double* d_vec = nullptr;   // m-vector scratch of the Schur solves
  size_t cap_vec = 0;
  // cooperative solves: per-strip flags, the launch epoch, the failure flag
  // (device and a pinned host copy written after each solve)
  int* d_flags = nullptr;
  int cap_flags = 0;
  int epoch = 0;
  int* d_err = nullptr;
  int* h_err = nullptr;
  int coop_state = 0;  // 0 unknown, 1 cooperative launches available, 2 not
  int ensure_flags(int n);
  // epoch-tagged words of the pipelined solves (2 per value, zeroed when
  // allocated) and the largest block count known to fit co-resident
  unsigned long long* d_ll = nullptr;
  size_t cap_ll = 0;
  int ll_max_nb = -1, ll_max_nb32 = -1;
  unsigned ll_epoch = 0;  // last epoch used in d_ll (reset only with a fresh, zeroed d_ll)
  int ensure_ll(size_t nwords);
  // clear the give-up flag (device word, stream-ordered, and its pinned copy)
  // once a failure has been reported, so later solves start clean
  int clear_err(hipStream_t s);
  // single-launch panel factorization: epoch-tagged exchange words
  // (kPanelTagBlocks blocks of 8192, zeroed when allocated) and the last epoch
  unsigned long long* d_ptag = nullptr;
  unsigned ptag_epoch = 0;
  int ensure_panel_tags();
  // look-ahead Cholesky: side stream (default priority) and per-panel events
  hipStream_t side = nullptr;
  std::vector<hipEvent_t> evp, evn;
  hipEvent_t evjoin = nullptr;
  int side_waves = 0;  // grid bound of the side stream's updates
  int ensure_side(int npanels);
  // Optional timing of the factorization's two main kernels: HIP events
  // around each launch of the wide trailing update on the side stream
  // (evt*) and of the single-launch panel factorization on the main stream
  // (evq*), with the algorithmic flops of each launch.
  bool time_trailing = false;
  std::vector<hipEvent_t> evt0, evt1, evq0, evq1;
  std::vector<double> evt_flops, evq_flops;
  int trail_pending = 0, panel_pending = 0;
  double trail_ms = 0.0, trail_flops = 0.0, panel_ms = 0.0, panel_flops = 0.0;
  long trail_launches = 0, panel_launches = 0;
  int ensure_trailing_events(int n);
  int ensure_panel_events(int n);
  int collect_trailing();  // call once the factorization has completed
  int ensure_vec(size_t n);
  int ensure_dinv(size_t n);
  // Device copies of small descriptor arrays (MatDesc, GemmItem), kept
  // across calls and keyed by their bytes: an array equal to one uploaded
  // before is not copied again (the solver's loop passes the same ones every
  // iteration, and a copy from pageable memory holds the host until the
  // stream reaches it).  *out: the device copy.
  struct DescCopy {
    std::vector<unsigned char> bytes;
    void* d = nullptr;
  };
  std::vector<DescCopy> desc_pool;
  int desc_copy(hipStream_t s, const void* h, size_t bytes, const void** out);
  void release();
};

}  // namespace sdpa_amd
