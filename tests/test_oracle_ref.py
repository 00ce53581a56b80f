"""The oracle's restatements against the reference routines compiled from
the SDPA sources (oracle/_ref, built by `make -C oracle ref` where the
reference tree and MKL are present).  Skipped when that library is absent
(e.g. on a machine without the reference sources)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _spd(n, rng, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1, cond, n)) @ q.T


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 16, 33, 100])
def test_chol_adjust_matches_rATL_dpotrfL(n):
    rng = np.random.default_rng(n)
    A = _spd(n, rng)
    info_o, L_o = O.chol_adjust(A)
    flat = np.ascontiguousarray(A.reshape(-1, order="F"))
    info_r = O.ref().ref_chol_adjust(n, flat, n)
    L_r = flat.reshape(n, n, order="F")
    assert info_o == info_r == 0
    tril = np.tril_indices(n)
    np.testing.assert_allclose(L_o[tril], L_r[tril], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("n", [6, 9, 40])
def test_chol_adjust_singular_and_indefinite(n):
    rng = np.random.default_rng(100 + n)
    # exactly zero rows/columns: those pivots are exactly 0 and get the
    # 1e100 assignment (sdpa_dpotrf.cpp:31-33), whatever the blocking order
    A = _spd(n, rng)
    zero = rng.choice(n, max(1, n // 4), replace=False)
    A[zero, :] = 0.0
    A[:, zero] = 0.0
    info_o, L_o = O.chol_adjust(A)
    flat = np.ascontiguousarray(A.reshape(-1, order="F"))
    info_r = O.ref().ref_chol_adjust(n, flat, n)
    assert info_o == info_r == 0
    L_r = flat.reshape(n, n, order="F")
    assert (np.diag(L_o) > 1e40).sum() == len(zero)
    assert ((np.diag(L_o) > 1e40) == (np.diag(L_r) > 1e40)).all()
    tril = np.tril_indices(n)
    np.testing.assert_allclose(L_o[tril], L_r[tril], rtol=1e-10, atol=1e-12)
    # clearly indefinite: both fail at the same pivot
    B = _spd(n, rng) - 2e3 * np.eye(n)
    info_o, _ = O.chol_adjust(B)
    flat = np.ascontiguousarray(B.reshape(-1, order="F"))
    assert info_o == O.ref().ref_chol_adjust(n, flat, n) > 0


@pytest.mark.parametrize("n,seed", [(21, 0), (30, 1), (64, 2), (200, 3), (37, 4)])
def test_lanczos_matches_Lal_getMinEigen(n, seed):
    rng = np.random.default_rng(seed)
    X = _spd(n, rng, 1e2)
    D = rng.standard_normal((n, n)); D = (D + D.T) / 2
    _, Linv, _ = O.inv_chol(X)
    o = O.lanczos_min(Linv, D)
    r = O.ref().ref_lanczos_min(n, np.ascontiguousarray(Linv.reshape(-1, order="F")),
                                np.ascontiguousarray(D.reshape(-1, order="F")))
    assert o == pytest.approx(r, rel=1e-9, abs=1e-12)
    lam = np.linalg.eigvalsh(Linv @ D @ Linv.T).min()
    assert o <= lam + 1e-9 * abs(lam)  # the truncated estimate is a lower estimate


@pytest.mark.parametrize("n", [2, 5, 12, 20])
def test_small_min_eig_matches_dsyev_path(n):
    rng = np.random.default_rng(n)
    X = _spd(n, rng, 1e2)
    D = rng.standard_normal((n, n)); D = (D + D.T) / 2
    _, Linv, _ = O.inv_chol(X)
    T = Linv @ D @ Linv.T
    o = O.sym_min_eig(np.asfortranarray(T))
    r = O.ref().ref_small_min_eig(n, np.ascontiguousarray(Linv.reshape(-1, order="F")),
                                  np.ascontiguousarray(D.reshape(-1, order="F")))
    assert o == pytest.approx(r, rel=1e-10, abs=1e-12)


@pytest.mark.parametrize("n", [2, 3, 17, 64])  # size 1 is an LP block in SDPA
def test_inv_chol_matches_getInvCholAndInv(n):
    rng = np.random.default_rng(n)
    A = _spd(n, rng, 1e3)
    info, Linv, Ainv = O.inv_chol(A)
    Lr = np.zeros(n * n); Ar = np.zeros(n * n)
    ok = O.ref().ref_inv_chol(n, np.ascontiguousarray(A.reshape(-1, order="F")), Lr, Ar)
    assert info == 0 and ok == 1
    np.testing.assert_allclose(Linv, Lr.reshape(n, n, order="F"), rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(Ainv, Ar.reshape(n, n, order="F"), rtol=1e-10, atol=1e-14)
