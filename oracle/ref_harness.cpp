// ref_harness.cpp -- C entry points over routines of the SDPA reference sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) in the
// build container, against the reference sources where they lie, into
// oracle/_ref/.  It lets the tests pin the C restatement in
// oracle/sdpa_oracle.c against the reference's own arithmetic for the parts
// whose sources compile without MUMPS:
//   rATL_dpotrfL                (sdpa_dpotrf.cpp:218-265)
//   Lal::getMinEigen            (sdpa_linear.cpp:29-143, truncated Lanczos)
//   Lal::getMinEigenValue       (sdpa_linear.cpp:145-182, dsyev path)
//   Jal::getInvChol/getInvCholAndInv (sdpa_jordan.cpp:148-216)
// Only this file is ours; everything it calls is compiled from the reference.
#include <cstring>
#include <vector>
#include "sdpa_include.h"
#include "sdpa_jordan.h"
#include "sdpa_linear.h"
#include "sdpa_dpotrf.h"

using namespace sdpa;

// DenseMatrix(nRow, nCol, type) reads de_ele before setting it
// (sdpa_struct.cpp:748-782), so matrices here are default-constructed first.
static void dm_init(DenseMatrix& a, int n, const double* src) {
  a.initialize(n, n, DenseMatrix::DENSE);
  if (src) std::memcpy(a.de_ele, src, sizeof(double) * n * n);
}

extern "C" {

int ref_chol_adjust(int n, double* A, int lda) { return (int)rATL_dpotrfL(n, A, lda); }

double ref_lanczos_min(int n, const double* L, const double* D) {
  DenseMatrix l, x, Q;
  dm_init(l, n, L); dm_init(x, n, D); dm_init(Q, n, nullptr);
  Vector out(n), b(n), r(n), q(n), qold(n), w(n), tmp(n), d1(n), d2(n);
  Vector work(3 * n > 1 ? 3 * n - 1 : 1);
  return Lal::getMinEigen(l, x, Q, out, b, r, q, qold, w, tmp, d1, d2, work);
}

// lambda_min(L D L^T) by the n <= 20 branch of Jal::getMinEigen
// (sdpa_jordan.cpp:81-89): DLS2 = D * L^T, DLS1 = L * DLS2, then dsyev.
double ref_small_min_eig(int n, const double* L, const double* D) {
  DenseMatrix l, x, t1, t2;
  dm_init(l, n, L); dm_init(x, n, D); dm_init(t1, n, nullptr); dm_init(t2, n, nullptr);
  Lal::let(t2, '=', x, 'T', l);
  Lal::let(t1, '=', l, '*', t2);
  Vector ev(n), work(3 * n > 1 ? 3 * n - 1 : 1);
  return Lal::getMinEigenValue(t1, ev, work);
}

// Jal::getInvCholAndInv on a one-block space; returns 1 on success.
int ref_inv_chol(int n, const double* A, double* Linv, double* Ainv) {
  BlockStruct bs;
  bs.initialize(1);
  bs.blockStruct[0] = n;
  bs.blockType[0] = BlockStruct::btSDP;
  bs.makeInternalStructure();
  // DenseLinearSpace(bs) has the same uninitialized-pointer read, so
  // default-construct and initialize (sdpa_struct.cpp:1457-1481).
  DenseLinearSpace inv, invs, a, work;
  inv.initialize(bs); invs.initialize(bs); a.initialize(bs); work.initialize(bs);
  std::memcpy(a.SDP_block[0].de_ele, A, sizeof(double) * n * n);
  bool ok = Jal::getInvCholAndInv(inv, invs, a, work);
  if (ok) {
    std::memcpy(Linv, inv.SDP_block[0].de_ele, sizeof(double) * n * n);
    std::memcpy(Ainv, invs.SDP_block[0].de_ele, sizeof(double) * n * n);
  }
  return ok ? 1 : 0;
}

}  // extern "C"
