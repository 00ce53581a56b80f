// ctx.h -- the device-resident state behind one sdpa_gpu_ctx (internal).
#pragma once
#include "common.h"

#include <string>

namespace sdpa_amd {

// Constraint table of one SDP block (InputData::SDP_constraint and the
// SparseMatrix data of every A_k that has this block, sdpa_dataset.cpp:268-309).
struct BlockTable {
  int n = 0;          // block size
  long off = 0;       // packed offset
  int ncon = 0;       // constraints with data in this block
  int nF1 = 0, nF3 = 0;
  int nheavy = 0;
  int ndense = 0;     // constraints whose block is dense (changeToDense)
  int f3cols = 0;     // F3 by column blocks (nearly all constraints one entry, n <= 256)
  // device arrays, length ncon (or as noted)
  int* d_con = nullptr;    // global constraint index (increasing)
  int* d_estart = nullptr; // first entry in the entry pool
  int* d_ecnt = nullptr;   // stored entries (NonZeroCount)
  int* d_nzeff = nullptr;  // NonZeroEffect
  int* d_form = nullptr;   // 1, 2, 3 (sdpa_newton.cpp:515-695)
  int* d_dpos = nullptr;   // column in the dense pool of this block, -1 if sparse
  int* d_heavy = nullptr;  // local indices of heavy constraints (nheavy)
  int* d_f1own = nullptr;  // local indices of F1/F2 owners (nF1)
  int* d_dcon = nullptr;   // global constraint of each dense pool column (ndense)
  double* d_dense = nullptr;  // n*n x ndense dense pool (column per dense A_k)
  float* d_dense32 = nullptr; // the pool in FP32 (made by the first FP32 assembly)
  int* d_dlist = nullptr;     // local index of each dense pool column's constraint (ndense)
  // blocks with n <= 256: one 16-byte record per constraint (con, estart, nzeff,
  // min(ecnt, 9) | form << 4 | a << 8 | b << 16 with (a, b) its entry when it
  // has one) and that entry's value (0 otherwise)
  int* d_rmeta = nullptr;
  double* d_rval = nullptr;
  std::vector<int> h_con, h_form, h_dpos, h_ecnt, h_estart, h_f1own, h_nzeff;
  // F1 chunks: the owners per chunk and, per chunk, the first dense-pool
  // column that one of the chunk's owners owns (the Abar^T Gbar product
  // starts there); set at the block's first assembly
  int f1chunk = 0;
  std::vector<int> f1rlo;
};

// int4 words per visit record of the all-blocks F3 kernel
constexpr int kVisInt4 = 5;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  DenseWork dw;
  bool loaded = false;
  // problem structure
  int m = 0, nBlock = 0, nSDP = 0, nLP = 0, nDim = 0;
  long total = 0, lpOff = 0;
  std::vector<int> sdpN;
  std::vector<long> sdpOff;
  int* d_sdpN = nullptr;
  long* d_sdpOff = nullptr;
  std::vector<int> blockType, blockNumber, blockStruct;
  int nF[3] = {0, 0, 0};
  int schurDense = 1;
  bool bshifted = false;  // the sparse branch's diagonal shift is in d_B
  // sparse branch factored by tiles of the ordered matrix (when its analysis
  // finds the factor cheaper than the dense one by the reference's cost rule)
  int schurSparse = 0;
  TileSparse ts;
  bool bmatStoreFirst = false;  // single block covering all constraints, all F3
  // F3 of the SDP blocks of n <= kMultiN by one launch (several blocks,
  // m <= kF3MultiMaxM; larger blocks keep their per-block kernels): per
  // constraint j its visits, one per block where A_j has at most kLight
  // entries, in block order
  int f3multi = 0;
  int f3m_pmax = 0, f3m_nmax = 0;  // largest n (n + 1) / 2 and n of those blocks
  int* d_jptr = nullptr;  // m + 1
  // per visit kVisInt4 int4: the block's first entry record and chunks, n,
  // min(ecnt, 15) and (form == 3) of A_j, j's chunk; nzeff, estart, the
  // block's packed offset; a | b << 8 of A_j's first four entries (16 bits
  // each) and their values
  int* d_jent = nullptr;
  int* d_mrec = nullptr;  // every block's entry records (int4 each)
  std::vector<BlockTable> blk;
  // one allocation holding every block's small tables (d_con ... d_dlist,
  // d_rmeta, d_rval: the BlockTable pointers point into it; d_dense does not)
  void* d_blkarena = nullptr;
  // SDP sparse entry pool (local row/col, value)
  int* d_er = nullptr;
  int* d_ec = nullptr;
  double* d_ev = nullptr;
  long nent = 0;
  // LP coordinates: CSR over coordinates of (constraint, value)
  int* d_lpptr = nullptr;
  int* d_lpcon = nullptr;
  double* d_lpval = nullptr;
  // constraint-major sparse terms for A_i . D: (p1, p2, v), CSR over constraints
  int* d_cptr = nullptr;
  long* d_cp1 = nullptr;
  long* d_cp2 = nullptr;
  double* d_cv = nullptr;
  // position-major terms for sum_i y_i A_i: CSR over distinct packed positions
  long npos = 0;
  long* d_ppos = nullptr;
  long* d_pptr = nullptr;
  int* d_pcon = nullptr;
  double* d_pval = nullptr;
  // dense A blocks: (constraint, block) list for dots and GEMV
  int ndenseTot = 0;
  double* d_b = nullptr;   // m
  double* d_C = nullptr;   // packed dense C (internal, = -F0)
  // iterate and work (packed block space unless noted)
  double *d_X = nullptr, *d_Z = nullptr, *d_y = nullptr;
  double *d_invCholX = nullptr, *d_invCholZ = nullptr, *d_Zinv = nullptr;
  double *d_LX = nullptr, *d_LZ = nullptr;
  double *d_Dx = nullptr, *d_Dz = nullptr, *d_dy = nullptr;
  double *d_rzinv = nullptr, *d_xrdzinv = nullptr, *d_g = nullptr;
  double *d_w1 = nullptr, *d_w2 = nullptr, *d_w3 = nullptr;
  double *d_pvec = nullptr, *d_dmat = nullptr;  // residuals
  double *d_X0 = nullptr, *d_Z0 = nullptr;      // initial point (init-point runs)
  double* d_B = nullptr;   // m x m Schur matrix, column-major, lower triangle used
  double* d_Binv = nullptr;  // inverses of the 64x64 diagonal blocks of its factor
  double* d_Winv = nullptr;  // inverses of its 256x256 diagonal blocks (for the solves)
  double* d_G = nullptr;   // F1 scratch: n*n per owner in a chunk
  size_t capG = 0;
  double* d_F1C = nullptr;  // F1 dense-pair products
  size_t capF1C = 0;
  double* d_F1P = nullptr;  // their partial sums over slices of the n^2 dimension
  double* d_hpart = nullptr;  // F3 heavy-pair partial sums (nheavy^2 x slices)
  size_t capHpart = 0;
  size_t capF1P = 0;
  float* d_xz32 = nullptr;  // FP32 assembly: one block's X and Z^-1 in FP32
  size_t capXZ32 = 0;
  int* d_info = nullptr;   // factorization info (one per batch item)
  double* d_red = nullptr;  // reduction partials (kMaxRed x 4096), then the results (64)
  double* h_red = nullptr;  // pinned host mirror of reduction results
  double* h_eig = nullptr;  // pinned: the step-length estimates per SDP block (2 nSDP)
  // pinned staging ring of the loader's dense-pool upload (kStageBufs
  // buffers of kStageBytes; each one's last copy recorded in ev_stage)
  void* h_stage[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_stage[4] = {nullptr, nullptr, nullptr, nullptr};
  int* h_info = nullptr;    // pinned
  double* d_eig = nullptr;  // per-item eigen results
  // mixed precision (made on first use): the FP32 factor of S B S, the
  // scaling, the FP64 panel and diagonal-block buffers, refinement vectors
  float* d_B32 = nullptr;
  double *d_scale = nullptr, *d_P64 = nullptr, *d_D64 = nullptr;
  double *d_mr = nullptr, *d_mx = nullptr, *d_md = nullptr;
  double* d_symv = nullptr;   // symmetric-product partials, nt * nt * 64
  double* d_mnorm = nullptr;  // norms of the equilibrated residual, solution, rhs
  double* h_mnorm = nullptr;  // pinned mirror
  int* d_mflags = nullptr;    // [0] failing column, [1] non-positive diagonal
  int* h_mflags = nullptr;    // pinned mirror
  unsigned long long* d_bmax = nullptr;
  int mixed_m = 0;            // the size the buffers were made for
  int factor_mixed = 0;       // the current factor is the mixed one (B intact)
  double mx_anorm = 0.0;      // ||S B S||_inf of that factorization
  long mx_factors = 0, mx_fallbacks = 0, mx_solves = 0, mx_refine = 0, mx_stalls = 0;
  // timing (ms), accumulated since the last reset
  hipEvent_t ev[8];
  double t_bmat = 0, t_chol = 0, t_solve = 0, t_eig = 0, t_inv = 0;
  // finer buckets (ms): the F1/F2 share of bMat, gVec, direction,
  // residuals, update, inner products
  double t_f1 = 0, t_gvec = 0, t_dir = 0, t_res = 0, t_upd = 0, t_dots = 0;
  int n_bmat = 0, n_chol = 0, n_solve = 0;
  bool timing = false;
  std::string last_error;
};

// ops.hip: reductions per batch (dots, norms)
constexpr int kMaxRed = 16;
// schur.hip
constexpr int kF3MultiMaxM = 8192;  // the all-blocks F3 kernel keeps a Schur column in LDS
constexpr int kMultiN = 64;         // ... and takes the blocks up to this size
int make_bmat(Ctx* c, int precision);  // 1: F1/F2 products in FP32
int f3multi_prepare(int m, int pmax, int nmax);  // LDS size of that kernel
size_t f3multi_lds(int m, int pmax, int nmax, int* soff);  // its LDS in doubles
// ops.hip
enum DotCode {
  DOT_CX = 0, DOT_BY, DOT_CDX, DOT_BDY, DOT_XDZ, DOT_DXZ, DOT_DXDZ, DOT_XZ,
  DOT_TRX, DOT_TRZ, DOT_X0Z, DOT_XZ0, DOT_COUNT
};
int compute_dots(Ctx* c, int nq, const int* codes, double* out);
int residuals(Ctx* c, double* normP, double* normD);
int make_gvec(Ctx* c, int corrector, double target, int infeasible);
int make_direction(Ctx* c, int infeasible);
int update_point(Ctx* c, double alphaP, double alphaD);
int inverse_chol(Ctx* c);
int sum_yA(Ctx* c, const double* d_yv, double* d_out);  // out = sum_i y_i A_i (packed)
int a_dot(Ctx* c, const double* d_D, double* d_out);    // out_i = A_i . D
int triple_product(Ctx* c, const double* A, const double* B, const double* C, double* R);
// launches whose scalar results are read after one later synchronization:
// result slots of d_red's result area (kRedResults doubles, mirrored in h_red)
constexpr int kRedResults = 64;
constexpr int kSlotResid = 40;  // residual max norms (2)
constexpr int kSlotLpMin = 44;  // LP minima of the step-length estimates (2)
int dots_enqueue(Ctx* c, int nq, const int* codes, int slot0);
int residuals_enqueue(Ctx* c, int slot);
int lp_min_enqueue(Ctx* c, int slot);
int inverse_chol_enqueue(Ctx* c);
int inverse_chol_status(const Ctx* c);
int min_eig_enqueue(Ctx* c, int lp_slot);
void min_eig_finish(const Ctx* c, int lp_slot, double* lamX, double* lamZ);
// eig.hip
int min_eig(Ctx* c, double* lamX, double* lamZ);
// gemm.hip: C_b = alpha op(A_b) op(B_b) + beta C_b for b < batch (FP64 MFMA,
// column-major; t = 1 transposes; s* are the batch strides in elements, 0
// for an operand shared by the batch).  Single blocks of at least kBlasMinN
// rows take these products; many small blocks stay in one batched launch of
// the small-tile kernel.
constexpr int kBlasMinN = 128;
int dgemm(hipStream_t s, int ta, int tb, int M, int N, int K, double alpha, const double* A,
          int lda, long sA, const double* B, int ldb, long sB, double beta, double* C, int ldc,
          long sC, int batch);
// the same in FP32 (v_mfma_f32_16x16x4f32, FP32 accumulation)
int sgemm(hipStream_t s, int ta, int tb, int M, int N, int K, float alpha, const float* A,
          int lda, long sA, const float* B, int ldb, long sB, float beta, float* C, int ldc,
          long sC, int batch);
// mixed.hip
int factor_mixed(Ctx* c);  // 0 ok; otherwise the FP64 path must be used
int solve_mixed(Ctx* c);   // 0 converged (dy set); 2 refinement stalled
void free_mixed(Ctx* c);

}  // namespace sdpa_amd
