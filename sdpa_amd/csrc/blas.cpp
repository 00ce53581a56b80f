// blas.cpp -- plain library GEMMs through rocBLAS (FP64, MFMA kernels of the
// library): the F1/F2 products of dense constraint blocks (X A_j, then
// (X A_j) Z^-1 for a chunk of constraints at once, and Abar^T Gbar), where
// the operands are whole matrices with no structure to fuse.
#include <rocblas/rocblas.h>

#include "ctx.h"

namespace sdpa_amd {

static rocblas_operation op(int t) { return t ? rocblas_operation_transpose : rocblas_operation_none; }

static int handle(Ctx* c, rocblas_handle* h) {
  if (!c->blas) {
    rocblas_handle nh = nullptr;
    if (rocblas_create_handle(&nh) != rocblas_status_success) {
      c->last_error = "rocblas_create_handle failed";
      return -41;
    }
    c->blas = nh;
  }
  *h = static_cast<rocblas_handle>(c->blas);
  if (rocblas_set_stream(*h, c->stream) != rocblas_status_success) {
    c->last_error = "rocblas_set_stream failed";
    return -41;
  }
  return 0;
}

void blas_release(Ctx* c) {
  if (c->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
  c->blas = nullptr;
}

int blas_gemm_strided(Ctx* c, int ta, int tb, int M, int N, int K, double alpha, const double* A,
                      int lda, long sA, const double* B, int ldb, long sB, double beta, double* C,
                      int ldc, long sC, int batch) {
  rocblas_handle h;
  if (int e = handle(c, &h)) return e;
  const rocblas_status st = rocblas_dgemm_strided_batched(h, op(ta), op(tb), M, N, K, &alpha, A, lda,
                                                          sA, B, ldb, sB, &beta, C, ldc, sC, batch);
  if (st != rocblas_status_success) {
    c->last_error = "rocblas_dgemm_strided_batched failed";
    return -42;
  }
  return 0;
}

int blas_gemm(Ctx* c, int ta, int tb, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc) {
  rocblas_handle h;
  if (int e = handle(c, &h)) return e;
  const rocblas_status st =
      rocblas_dgemm(h, op(ta), op(tb), M, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc);
  if (st != rocblas_status_success) {
    c->last_error = "rocblas_dgemm failed";
    return -42;
  }
  return 0;
}

}  // namespace sdpa_amd
