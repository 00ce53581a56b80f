"""The oracle's bMat, eigen and factor routines against what the mathematics
fixes: the textbook Schur complement B_ij = Tr(A_i X A_j Z^-1) evaluated by
brute force on dense matrices, the MaxCut and Lovasz closed forms, LAPACK-
equivalent library routines (numpy) on SPD inputs, and the formula-choice
outcomes stated for the benchmark families (MaxCut/Lovasz all F3, dense
data all F1/F2)."""
import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import problems as PR


def dense_mats(d):
    """A_1..A_m and C as dense per-SDP-block matrices (internal convention)."""
    sdp = [s for s in d.block_struct if s >= 2]
    blocks = {}
    idx = 0
    for l, s in enumerate(d.block_struct):
        if s >= 2:
            blocks[l + 1] = (idx, s)
            idx += 1
    A = [[np.zeros((s, s)) for s in sdp] for _ in range(d.m + 1)]
    for r in range(d.nnz):
        k, l, i, j, v = int(d.k[r]), int(d.l[r]), int(d.i[r]) - 1, int(d.j[r]) - 1, d.v[r]
        if l not in blocks:
            continue
        q, s = blocks[l]
        vv = -v if k == 0 else v
        A[k][q][i, j] += vv
        if i != j:
            A[k][q][j, i] += vv
    return A


def random_spd(n, rng, cond=50.0):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1, cond, n)) @ q.T


def packed(mats):
    return np.concatenate([M.reshape(-1, order="F") for M in mats])


def brute_bmat(d, Xs, Zis):
    A = dense_mats(d)
    m = d.m
    B = np.zeros((m, m))
    for q in range(len(Xs)):
        XA = [Xs[q] @ A[k + 1][q] @ Zis[q] for k in range(m)]
        for i in range(m):
            for j in range(m):
                B[i, j] += np.sum(A[i + 1][q] * XA[j].T)
    return B


def sparse_random(n, m, nnz_per, seed):
    rng = np.random.default_rng(seed)
    recs = []
    for k in range(1, m + 1):
        cnt = int(rng.integers(1, nnz_per + 1))
        seen = set()
        for _ in range(cnt):
            a, b = sorted(rng.integers(0, n, 2))
            if (a, b) in seen:
                continue
            seen.add((a, b))
            recs.append((k, 1, a + 1, b + 1, rng.standard_normal()))
    for a in range(n):
        recs.append((0, 1, a + 1, a + 1, 1.0))
    K, L, I, J, V = PR._records(recs)
    return PR._make(m, [n], np.ones(m), K, L, I, J, V, f"sparse{n}_{m}_{seed}")


def test_maxcut_bmat_is_hadamard():
    d = PR.maxcut(30, 0.3, seed=7)
    P = O.Problem(d)
    assert (P.formula_counts() == [0, 0, d.m]).all()
    rng = np.random.default_rng(0)
    X, Z = random_spd(30, rng), random_spd(30, rng)
    Zi = np.linalg.inv(Z)
    B = P.make_bmat(packed([X]), packed([Zi]))
    np.testing.assert_allclose(B, X * Zi, rtol=1e-13, atol=1e-15)


def test_lovasz_bmat_closed_form():
    d = PR.lovasz(14, 0.4, seed=3)
    P = O.Problem(d)
    rng = np.random.default_rng(1)
    n = 14
    X, Z = random_spd(n, rng), random_spd(n, rng)
    Zi = np.linalg.inv(Z)
    B = P.make_bmat(packed([X]), packed([Zi]))
    kk = d.k
    edges = [(int(d.i[r]) - 1, int(d.j[r]) - 1) for r in range(d.nnz) if kk[r] >= 2]
    XZ = X @ Zi
    # trace row: B_00 = sum X o Zi ; B_0,cd = (X Zi)_cd + (X Zi)_dc
    assert B[0, 0] == pytest.approx(np.sum(X * Zi), rel=1e-12)
    for e, (c, dd) in enumerate(edges):
        assert B[e + 1, 0] == pytest.approx(XZ[c, dd] + XZ[dd, c], rel=1e-11, abs=1e-13)
    for e1, (a, b) in enumerate(edges):
        for e2, (c, dd) in enumerate(edges):
            want = X[b, c] * Zi[a, dd] + X[b, dd] * Zi[a, c] + X[a, c] * Zi[b, dd] + X[a, dd] * Zi[b, c]
            assert B[e1 + 1, e2 + 1] == pytest.approx(want, rel=1e-11, abs=1e-13)


@pytest.mark.parametrize("kind", ["dense", "sparse_mixed", "multiblock"])
def test_bmat_brute_force_trace(kind):
    rng = np.random.default_rng(11)
    if kind == "dense":
        d = PR.dense_random(n=7, m=6, seed=2)
    elif kind == "sparse_mixed":
        d = sparse_random(30, 40, 120, seed=5)
    else:
        d = PR.multi_block(n_sdp_blocks=3, nmin=3, nmax=6, lp=4, m=7, density=0.3, seed=9)
    P = O.Problem(d)
    cnt = P.formula_counts()
    if kind == "dense":
        assert cnt[2] == 0 and cnt[0] + cnt[1] == d.m
    if kind == "sparse_mixed":
        assert (cnt > 0).sum() >= 2, cnt  # at least two different formulas in play
    sizes = [int(s) for s in d.block_struct if s >= 2]
    Xs = [random_spd(n, rng) for n in sizes]
    Zs = [random_spd(n, rng) for n in sizes]
    Zis = [np.linalg.inv(Z) for Z in Zs]
    nlp = sum(-int(s) for s in d.block_struct if s < 0)
    xl = rng.uniform(1, 2, nlp); zl = rng.uniform(1, 2, nlp)
    B = P.make_bmat(np.concatenate([packed(Xs), xl]), np.concatenate([packed(Zis), 1 / zl]))
    want = brute_bmat(d, Xs, Zis) if sizes else np.zeros((d.m, d.m))
    if nlp:
        Alp = np.zeros((d.m, nlp))
        off = {}
        t = 0
        for l, s in enumerate(d.block_struct):
            if s < 0:
                off[l + 1] = t
                t += -s
        for r in range(d.nnz):
            if int(d.k[r]) > 0 and int(d.l[r]) in off:
                Alp[int(d.k[r]) - 1, off[int(d.l[r])] + int(d.i[r]) - 1] += d.v[r]
        want += (Alp * (xl / zl)) @ Alp.T
    np.testing.assert_allclose(B, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    np.testing.assert_allclose(B, B.T, rtol=0, atol=1e-12 * np.abs(B).max())


def test_tridiag_ql_matches_eigh():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 17, 40):
        d = rng.standard_normal(n); e = rng.standard_normal(max(n - 1, 0))
        info, w, Z = O.tridiag_ql(d, e)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert info == 0
        np.testing.assert_allclose(w, np.linalg.eigvalsh(T), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(T @ Z, Z * w, atol=1e-11)
        np.testing.assert_allclose(Z.T @ Z, np.eye(n), atol=1e-12)


def test_jacobi_min_eig_matches_eigvalsh():
    rng = np.random.default_rng(4)
    for n in (2, 3, 8, 20):
        A = rng.standard_normal((n, n)); A = A + A.T
        assert O.sym_min_eig(np.asfortranarray(A)) == pytest.approx(np.linalg.eigvalsh(A)[0], rel=1e-12, abs=1e-12)


def test_chol_solve_matches_numpy():
    rng = np.random.default_rng(5)
    A = random_spd(50, rng, 1e4)
    g = rng.standard_normal(50)
    info, L = O.chol_adjust(A)
    assert info == 0
    np.testing.assert_allclose(np.tril(L) @ np.tril(L).T, A, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(O.chol_solve(L, g), np.linalg.solve(A, g), rtol=1e-9)


def test_lanczos_is_lower_estimate_and_tight_on_easy_spectra():
    """min - |err*beta| bounds lambda_min from below (Kaniel-Paige); on a
    spectrum with an isolated minimum it converges to it."""
    rng = np.random.default_rng(6)
    n = 60
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.linspace(1.0, 2.0, n); ev[0] = -3.0
    D = (q * ev) @ q.T
    L = np.eye(n)
    est = O.lanczos_min(L, D)
    assert est <= -3.0 + 1e-12
    assert est == pytest.approx(-3.0, abs=5e-3)


def test_maxcut_solution_bound_and_schur_dense():
    d = PR.maxcut(40, 0.2, seed=2)
    P = O.Problem(d)
    assert P.schur_dense
    out = P.solve()
    assert out.phase == "pdOPT"
    # SDP value <= (n/4) lambda_max(L); the MaxCut relaxation's optimum
    # c.x = sum x_i equals max (L/4).Y over diag(Y) = 1
    A = dense_mats(d)
    Lap4 = -A[0][0]
    assert out.dual_obj <= d.m * np.linalg.eigvalsh(Lap4)[-1] + 1e-6
    assert out.primal_obj == pytest.approx(out.dual_obj, rel=1e-6)


def test_formula_choice_on_benchmark_families():
    """MaxCut n=100 and Lovasz n=200 select F3 for every constraint (sparse
    data, n <= 200), the dense family only F1/F2 (sdpa_newton.cpp:631-676)."""
    for d in (PR.maxcut(100, 0.1, 1), PR.lovasz(200, 0.3, 2)):
        P = O.Problem(d)
        assert (P.formula_counts() == [0, 0, d.m]).all(), d.name
        assert P.schur_dense
    P = O.Problem(PR.dense_random(n=20, m=30, seed=3))
    c = P.formula_counts()
    assert c[2] == 0


def _two_block_problem(ranges, m=120, n=4):
    """Two SDP blocks; constraint k (1-based) touches block b when k-1 is in
    ranges[b], with one diagonal entry there."""
    from sdpa_amd.problems import SdpData
    K, L, I, J, V = [], [], [], [], []
    for b, (lo, hi) in enumerate(ranges):
        for k in range(lo, hi):
            K.append(k + 1); L.append(b + 1); I.append(1 + k % n); J.append(1 + k % n); V.append(1.0)
    for b in range(2):  # F0 = identity on both blocks
        for t in range(n):
            K.append(0); L.append(b + 1); I.append(t + 1); J.append(t + 1); V.append(1.0)
    arr = lambda x, t: np.asarray(x, t)
    return SdpData(m, arr([n, n], np.int32), np.ones(m), arr(K, np.int32), arr(L, np.int32),
                   arr(I, np.int32), arr(J, np.int32), arr(V, np.float64))


def test_schur_dense_aggregate_pattern_rule():
    """sdpa_chordal.cpp:404-410: with m > 100 and no block above m*sqrt(0.7)
    constraints, the Schur path is dense iff the aggregate pattern (pairs of
    constraints sharing a block, plus the diagonal) exceeds 0.7 m^2.
    Overlapping 90-constraint sets: 2*90^2 - 60^2 = 12600 > 10080 -> dense;
    disjoint 60-constraint sets: 2*60^2 = 7200 -> not dense."""
    assert O.Problem(_two_block_problem([(0, 90), (30, 120)])).schur_dense
    assert not O.Problem(_two_block_problem([(0, 60), (60, 120)])).schur_dense
