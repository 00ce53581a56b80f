"""The oracle's Newton direction (orc_direction: Make_gVec, the Schur solve,
compute_DzMat / compute_DxMat, sdpa_newton.cpp:793-883, 2672-2752) against
the equations that define the HKM search direction (Helmberg-Rendl-
Vanderbei-Wolkowicz, Kojima-Shindoh-Hara, Monteiro), checked with numpy on
dense matrices:

    A_i . Dx = p_i                          (p = b - A(X), primal residual)
    Dz + sum_i dy_i A_i = [infeasible] D    (D = C - sum_i y_i A_i - Z)
    Dx = sym(mu' Z^-1 - X - [corrector] Dx' Dz' Z^-1 - X Dz Z^-1)

per SDP block (and the same scalar relations on the LP coordinates), where
mu' is the step's target and Dx', Dz' the predictor's direction.  The three
together fix the direction; a dropped term or a wrong sign in the right-hand
side, the Schur system or the back-substitution breaks one of them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import problems as PR

from test_oracle_math import sparse_random


def packed_mats(d):
    """Rows: C (= -F0) then A_1..A_m, each in the packed block space with the
    SDP blocks full and symmetric."""
    offs, off = [], 0
    for s in d.block_struct:
        s = int(s)
        offs.append(off)
        off += s * s if s >= 2 else abs(s)
    M = np.zeros((d.m + 1, off))
    for r in range(d.nnz):
        k, l, i, j, v = int(d.k[r]), int(d.l[r]) - 1, int(d.i[r]) - 1, int(d.j[r]) - 1, d.v[r]
        s = int(d.block_struct[l])
        vv = -v if k == 0 else v
        if s >= 2:
            M[k, offs[l] + i + s * j] += vv
            if i != j:
                M[k, offs[l] + j + s * i] += vv
        else:
            M[k, offs[l] + i] += vv
    return M, offs


def random_iterate(d, rng):
    X, Z = [], []
    for s in d.block_struct:
        s = int(s)
        if s >= 2:
            for out in (X, Z):
                q, _ = np.linalg.qr(rng.standard_normal((s, s)))
                out.append(((q * rng.uniform(0.5, 3.0, s)) @ q.T).reshape(-1, order="F"))
        else:
            X.append(rng.uniform(0.5, 3.0, abs(s)))
            Z.append(rng.uniform(0.5, 3.0, abs(s)))
    return np.concatenate(X), rng.standard_normal(d.m), np.concatenate(Z)


def check_direction(d, X, y, Z, infeasible, mu, bp, bc, tol=1e-9):
    P = O.Problem(d)
    pred, corr = P.direction(X, y, Z, infeasible, mu, bp, bc)
    M, offs = packed_mats(d)
    C, A = M[0], M[1:]
    p = d.c - A @ X
    D = C - A.T @ y - Z
    scale = 1.0 + np.abs(X).max() + np.abs(Z).max()
    for step, beta, prev in ((pred, bp, None), (corr, bc, pred)):
        Dx, Dz, dy = step["Dx"], step["Dz"], step["dy"]
        np.testing.assert_allclose(A @ Dx, p, rtol=tol, atol=tol * scale * (1 + np.abs(Dx).max()))
        np.testing.assert_allclose(Dz + A.T @ dy, D if infeasible else 0.0 * D, rtol=tol,
                                   atol=tol * scale * (1 + np.abs(dy).max()))
        target = beta * mu
        for l, s in enumerate(d.block_struct):
            s = int(s)
            o = offs[l]
            if s >= 2:
                sl = slice(o, o + s * s)
                Xb, Zb = X[sl].reshape(s, s, order="F"), Z[sl].reshape(s, s, order="F")
                Zi = np.linalg.inv(Zb)
                R = target * Zi - Xb - Xb @ Dz[sl].reshape(s, s, order="F") @ Zi
                if prev is not None:
                    R -= prev["Dx"][sl].reshape(s, s, order="F") @ \
                        prev["Dz"][sl].reshape(s, s, order="F") @ Zi
                np.testing.assert_allclose(Dx[sl].reshape(s, s, order="F"), (R + R.T) / 2,
                                           rtol=tol, atol=tol * scale * (1 + np.abs(R).max()))
            else:
                sl = slice(o, o + abs(s))
                R = target / Z[sl] - X[sl] - X[sl] * Dz[sl] / Z[sl]
                if prev is not None:
                    R -= prev["Dx"][sl] * prev["Dz"][sl] / Z[sl]
                np.testing.assert_allclose(Dx[sl], R, rtol=tol, atol=tol * scale)
    return pred, corr


CASES = {
    "sparse_F3": lambda: sparse_random(12, 20, 4, seed=3),
    "dense_F1": lambda: PR.dense_random(n=8, m=12, seed=4),
    "lovasz": lambda: PR.lovasz(14, 0.3, seed=5),
    "multiblock_lp": lambda: PR.multi_block(n_sdp_blocks=4, nmin=3, nmax=9, lp=6, m=15,
                                            density=0.1, seed=6),
}


@pytest.mark.parametrize("infeasible", [True, False])
@pytest.mark.parametrize("name", list(CASES))
def test_direction_satisfies_the_newton_equations(name, infeasible):
    d = CASES[name]()
    rng = np.random.default_rng(len(name))
    X, y, Z = random_iterate(d, rng)
    check_direction(d, X, y, Z, infeasible, mu=0.7, bp=0.2, bc=0.35)


def test_direction_at_the_identity_start_is_the_solvers_first_step():
    # from X = Z = lambda I, y = 0 the oracle's solve takes this direction in
    # its first predictor: the first trace line's objectives move by it
    d = PR.maxcut(10, 0.4, seed=2)
    lam = 100.0
    P = O.Problem(d)
    T = P.total
    X = np.zeros(T); Z = np.zeros(T)
    n = int(d.block_struct[0])
    X[::n + 1] = lam; Z[::n + 1] = lam
    y = np.zeros(d.m)
    pred, _ = check_direction(d, X, y, Z, True, lam * lam, 0.2, 0.2)
    # the direction from the start makes A(X + Dx) = b exactly
    M, _ = packed_mats(d)
    np.testing.assert_allclose(M[1:] @ (X + pred["Dx"]), d.c, rtol=1e-10, atol=1e-9)


def test_sparse_schur_decision_solves_the_shifted_system():
    # where Chordal::ordering_bMat leaves the dense path the reference factors
    # bMat + 1e-10 I (sdpa_chordal.cpp:25, 526-534): the direction then
    # satisfies A(Dx) = p - 1e-10 dy instead of A(Dx) = p (the dense path's
    # equations hold to the rounding level on the same data)
    for per_block, dense in ((60, False), (50, True)):
        # two blocks: m = 120 (sparse decision) or m = 100 (m <= 100: dense)
        d = PR.block_separated(per_block=per_block)
        P = O.Problem(d)
        assert P.schur_dense == dense
        rng = np.random.default_rng(7)
        X, y, Z = random_iterate(d, rng)
        pred, _ = P.direction(X, y, Z, True, 0.5, 0.2, 0.1)
        M, _ = packed_mats(d)
        r = M[1:] @ pred["Dx"] - (d.c - M[1:] @ X)
        want = np.zeros_like(r) if dense else -1.0e-10 * pred["dy"]
        np.testing.assert_allclose(r, want, rtol=0, atol=1e-14 * (1 + np.abs(pred["dy"]).max()))
