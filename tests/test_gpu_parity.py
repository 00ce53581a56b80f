"""HIP path (through the C ABI of include/sdpa_gpu.h and the solver wrapper)
against the CPU oracle on the same seeded inputs.

Tolerances: every kernel is FP64.  Per-element results that the oracle
computes with a different summation order are compared at rtol 1e-11
(bMat, factors, inverses); the solve at 1e-9 relative to the solution norm;
the truncated Lanczos estimate at 1e-7 relative (the device applies
L D L^T as one explicit matrix, the reference applies L, D, L^T in turn);
whole solves: same phase, iteration counts within one, objectives within
1e-6 relative and gap/feasibility at the 1e-7 stopping level.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import problems as PR

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sdpa_amd")


def spd(n, rng, cond=30.0):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1, cond, n)) @ q.T


def packed_random(d, rng, cond=30.0):
    """Random SPD X, Z per SDP block and positive LP entries, packed."""
    Xs, Zs, Zis = [], [], []
    for s in d.block_struct:
        s = int(s)
        if s >= 2:
            X, Z = spd(s, rng, cond), spd(s, rng, cond)
            Xs.append(X.reshape(-1, order="F")); Zs.append(Z.reshape(-1, order="F"))
            Zis.append(np.linalg.inv(Z).reshape(-1, order="F"))
    nlp = sum(-int(s) if s < 0 else (1 if s == 1 else 0) for s in d.block_struct)
    xl, zl = rng.uniform(0.5, 2, nlp), rng.uniform(0.5, 2, nlp)
    X = np.concatenate(Xs + [xl]); Z = np.concatenate(Zs + [zl]); Zi = np.concatenate(Zis + [1 / zl])
    return X, Z, Zi


def ctx_for(d):
    return sa.GpuContext(0).load(d)


def sparse_mixed(n=30, m=60, nnz_per=60, seed=5):
    """Sparse constraints with 1..nnz_per entries: F1, F2 and F3 owners, many
    heavy lists (more than 8 entries) and heavy-heavy pairs."""
    rng = np.random.default_rng(seed)
    recs = []
    for k in range(1, m + 1):
        seen = set()
        for _ in range(int(rng.integers(1, nnz_per + 1))):
            a, b = sorted(rng.integers(0, n, 2))
            if (a, b) not in seen:
                seen.add((a, b))
                recs.append((k, 1, a + 1, b + 1, rng.standard_normal()))
    for a in range(n):
        recs.append((0, 1, a + 1, a + 1, 1.0))
    K, L, I, J, V = PR._records(recs)
    return PR._make(m, [n], np.ones(m), K, L, I, J, V, f"sparse_mixed{n}_{m}")


CASES = {
    "maxcut100": lambda: PR.maxcut(100, 0.1, 1),
    "maxcut60_trace": lambda: PR.maxcut(60, 0.2, 4, trace_row=True),
    "sparse_mixed": sparse_mixed,
    "lovasz40": lambda: PR.lovasz(40, 0.3, 5),
    # one-entry constraints with a heavy trace row: the column-block F3 kernel
    # (with the heavy-pair slices) on a 541- and an 81-constraint block
    "lovasz60": lambda: PR.lovasz(60, 0.3, 7),
    "maxcut80_trace": lambda: PR.maxcut(80, 0.2, 8, trace_row=True),
    "dense_n20_m30": lambda: PR.dense_random(n=20, m=30, seed=2),
    "dense_n70_m40": lambda: PR.dense_random(n=70, m=40, seed=6),
    "multiblock_lp": lambda: PR.multi_block(n_sdp_blocks=6, nmin=3, nmax=30, lp=12, m=40,
                                            density=0.05, seed=9),
    # constraints spread over 12 blocks, with heavy lists (> 8 entries in a
    # block) and dense F1 owners (small blocks) added after the all-blocks F3
    "multiblock_wide": lambda: PR.multi_block(n_sdp_blocks=12, nmin=3, nmax=40, lp=30, m=300,
                                              density=0.02, seed=11),
    # blocks above 64 (per-block kernels adding to B) beside small ones
    "multiblock_sizes": lambda: PR.multi_block(n_sdp_blocks=6, nmin=10, nmax=120, lp=10, m=150,
                                               density=0.01, seed=13),
    # small blocks (n 8..12) whose constraints have 1..10 entries each: the
    # all-blocks F3 kernel's entry chunks padded at constraint boundaries, its
    # column entries beyond the four staged ones, and heavy lists beside them
    "multiblock_entries": lambda: PR.multi_block(n_sdp_blocks=10, nmin=8, nmax=12, lp=0, m=400,
                                                 density=0.05, seed=21),
    # the benchmark workloads at full size: Lovasz theta G(200, 0.3), m = 6011
    # (two 4096-row segments of the column-block F3 kernel plus the trace
    # row), and the 200-block problem, m = 5000 (the all-blocks F3 kernel with
    # lists of more than 512 constraints per block, the LP term, dense F1
    # owners in the small blocks)
    "lovasz_m6011": lambda: PR.lovasz(200, 0.3, 2),
    "multiblock_m5000": lambda: PR.multi_block(),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bmat_matches_oracle(name):
    d = CASES[name]()
    rng = np.random.default_rng(1)
    X, Z, Zi = packed_random(d, rng)
    g = ctx_for(d)
    g.set("X", X); g.set("Zinv", Zi)
    g.make_bmat()
    B = g.bmat()
    P = O.Problem(d)
    want = P.make_bmat(X, Zi)
    tril = np.tril_indices(d.m)
    scale = np.abs(want).max()
    np.testing.assert_allclose(B[tril], want[tril], rtol=1e-11, atol=1e-13 * scale)
    # changeToDense and computeFormula_SDP: the same F1/F2/F3 choice per
    # (constraint, block) pair, and the same dense/sparse Schur decision
    assert list(g.formula_counts()) == list(P.formula_counts()), name
    assert bool(g.info[9]) == P.schur_dense, name
    print(name, "formulas", g.formula_counts())


# 1300 and 2100: the look-ahead schedule's side-stream launches (m > 768) and
# the wide-tile trailing update (more than 256 columns left, m > 1024)
@pytest.mark.parametrize("m", [2, 5, 64, 65, 130, 300, 700, 1300, 2100])
def test_factor_and_solve_match_oracle(m):
    rng = np.random.default_rng(m)
    d = PR.maxcut(m, 0.5, seed=3)  # any problem with this m; only bMat is used
    g = ctx_for(d)
    A = spd(d.m, rng, 1e4)
    g.set("B", A.reshape(-1, order="F"))
    info = g.factor_bmat()
    assert info == 0
    L = g.bmat()
    info_o, Lo = O.chol_adjust(A)
    tril = np.tril_indices(d.m)
    np.testing.assert_allclose(L[tril], Lo[tril], rtol=1e-10, atol=1e-12)
    rhs = rng.standard_normal(d.m)
    g.set("g", rhs)
    g.solve_bmat()
    x = g.get("dy")
    want = np.linalg.solve(A, rhs)
    assert np.linalg.norm(x - want) <= 1e-9 * np.linalg.norm(want)


def test_factor_adjust_and_failure_rules():
    rng = np.random.default_rng(7)
    d = PR.maxcut(150, 0.5, seed=3)
    g = ctx_for(d)
    A = spd(150, rng, 1e3)
    zero = np.array([0, 63, 64, 100, 149])
    A[zero, :] = 0.0
    A[:, zero] = 0.0
    g.set("B", A.reshape(-1, order="F"))
    assert g.factor_bmat() == 0
    L = g.bmat()
    _, Lo = O.chol_adjust(A)
    big = np.diag(L) > 1e40
    assert (np.flatnonzero(big) == zero).all()
    tril = np.tril_indices(150)
    np.testing.assert_allclose(L[tril], Lo[tril], rtol=1e-10, atol=1e-12)
    B = spd(150, rng, 1e3) - 2e3 * np.eye(150)
    g.set("B", B.reshape(-1, order="F"))
    info_o, _ = O.chol_adjust(B)
    assert g.factor_bmat() == info_o > 0


@pytest.mark.parametrize("m", [1300, 2100])
def test_factor_adjust_and_failure_rules_single_launch_panels(m):
    # matrices beyond two 256-column panels take the single-launch panel
    # factorization: zero rows/columns (pivots below 1e-14 become 1e100) at
    # panel, step and block edges, and an indefinite matrix whose first
    # failing pivot lies in a later panel
    rng = np.random.default_rng(m)
    d = PR.maxcut(m, 0.5, seed=3)
    g = ctx_for(d)
    A = spd(m, rng, 1e3)
    zero = np.array([0, 63, 64, 255, 256, 300, 511, 700, m - 1])
    A[zero, :] = 0.0
    A[:, zero] = 0.0
    g.set("B", A.reshape(-1, order="F"))
    assert g.factor_bmat() == 0
    L = g.bmat()
    _, Lo = O.chol_adjust(A)
    big = np.diag(L) > 1e40
    assert (np.flatnonzero(big) == zero).all()
    tril = np.tril_indices(m)
    np.testing.assert_allclose(L[tril], Lo[tril], rtol=1e-10, atol=1e-12)
    # indefinite: a negative pivot appears past the first panels
    B = spd(m, rng, 1e3)
    B[900:, 900:] -= 5e3 * np.eye(m - 900)
    g.set("B", B.reshape(-1, order="F"))
    info_o, _ = O.chol_adjust(B)
    assert info_o > 900
    assert g.factor_bmat() == info_o


def test_inverse_cholesky_matches_oracle():
    d = PR.multi_block(n_sdp_blocks=5, nmin=2, nmax=140, lp=7, m=20, density=0.05, seed=4)
    rng = np.random.default_rng(2)
    X, Z, _ = packed_random(d, rng, cond=1e3)
    g = ctx_for(d)
    g.set("X", X); g.set("Z", Z)
    assert g.inv_chol() == 0
    icx, icz, zi = g.get("invCholX"), g.get("invCholZ"), g.get("Zinv")
    off = 0
    for s in d.block_struct:
        s = int(s)
        if s < 2:
            continue
        Xb = X[off:off + s * s].reshape(s, s, order="F")
        Zb = Z[off:off + s * s].reshape(s, s, order="F")
        _, Lx, _ = O.inv_chol(Xb)
        _, Lz, Zinv = O.inv_chol(Zb)
        np.testing.assert_allclose(icx[off:off + s * s].reshape(s, s, order="F"), Lx, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(icz[off:off + s * s].reshape(s, s, order="F"), Lz, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(zi[off:off + s * s].reshape(s, s, order="F"), Zinv, rtol=1e-9,
                                   atol=1e-12 * np.abs(Zinv).max())
        off += s * s
    np.testing.assert_allclose(icx[off:], 1 / np.sqrt(X[off:]), rtol=1e-14)
    np.testing.assert_allclose(zi[off:], 1 / Z[off:], rtol=1e-14)
    # a non-PD block is reported, as dpotrf's info != 0 is
    Xbad = X.copy()
    Xbad[0] = -1.0
    g.set("X", Xbad)
    assert g.inv_chol() == 1


def test_min_eig_matches_oracle():
    d = PR.multi_block(n_sdp_blocks=6, nmin=2, nmax=200, lp=5, m=20, density=0.05, seed=8)
    rng = np.random.default_rng(3)
    X, Z, _ = packed_random(d, rng, cond=50.0)
    g = ctx_for(d)
    g.set("X", X); g.set("Z", Z)
    assert g.inv_chol() == 0
    Dx = rng.standard_normal(len(X)); Dz = rng.standard_normal(len(X))
    # symmetric SDP blocks
    off = 0
    for s in d.block_struct:
        s = int(s)
        if s >= 2:
            for D in (Dx, Dz):
                b = D[off:off + s * s].reshape(s, s, order="F")
                D[off:off + s * s] = ((b + b.T) / 2).reshape(-1, order="F")
            off += s * s
    g.set("Dx", Dx); g.set("Dz", Dz)
    lx, lz = g.min_eig()
    P = O.Problem(d)
    ox = P.min_eig(g.get("invCholX"), Dx)
    oz = P.min_eig(g.get("invCholZ"), Dz)
    assert lx == pytest.approx(ox, rel=1e-7, abs=1e-10)
    assert lz == pytest.approx(oz, rel=1e-7, abs=1e-10)


@pytest.mark.parametrize("n", [2, 7, 20, 21, 100, 200, 208, 209, 300])
def test_min_eig_per_block_size(n):
    # one SDP block: up to 20 the tridiagonal reduction (the dsyev branch),
    # 21..208 the Lanczos kernel with the matrix held in registers, larger
    # blocks the one that streams it from memory
    d = PR.multi_block(n_sdp_blocks=1, nmin=n, nmax=n, lp=0, m=4, density=0.05, seed=n)
    rng = np.random.default_rng(n)
    X, Z, _ = packed_random(d, rng, cond=50.0)
    g = ctx_for(d)
    g.set("X", X); g.set("Z", Z)
    assert g.inv_chol() == 0
    D = []
    for _ in range(2):
        b = rng.standard_normal((n, n))
        D.append(((b + b.T) / 2).reshape(-1, order="F"))
    g.set("Dx", D[0]); g.set("Dz", D[1])
    lx, lz = g.min_eig()
    P = O.Problem(d)
    assert lx == pytest.approx(P.min_eig(g.get("invCholX"), D[0]), rel=1e-7, abs=1e-10)
    assert lz == pytest.approx(P.min_eig(g.get("invCholZ"), D[1]), rel=1e-7, abs=1e-10)


@pytest.mark.parametrize("name", ["example1", "example2"])
def test_solver_reproduces_manual(golden_dir, name):
    import json, os
    from sdpa_amd import dats
    gd = json.load(open(os.path.join(golden_dir, "manual_examples.json")))[name]
    s = sa.Solver(0).read_input(os.path.join(golden_dir, gd["file"]))
    r = s.solve()
    assert r.phase == gd["phase"]
    assert r.iterations == gd["iterations"]
    assert r.primal_obj == pytest.approx(gd["objValPrimal"], rel=1e-9)
    assert r.dual_obj == pytest.approx(gd["objValDual"], rel=1e-9)
    assert r.relative_gap == pytest.approx(gd["relative_gap"], rel=1e-5)
    o = O.Problem(dats.read(os.path.join(golden_dir, gd["file"]))).solve()
    for a, b in zip(r.trace, o.trace):
        for key, idx in (("mu", 0), ("objP", 3), ("objD", 4), ("beta", 7)):
            assert a[idx] == pytest.approx(b[key], rel=1e-7, abs=1e-9), (name, key)


SOLVE_CASES = {
    "maxcut100": lambda: PR.maxcut(100, 0.1, 1),
    "lovasz30": lambda: PR.lovasz(30, 0.3, 2),
    "dense_n25_m40": lambda: PR.dense_random(n=25, m=40, seed=3),
    "multiblock_lp": lambda: PR.multi_block(n_sdp_blocks=8, nmin=3, nmax=25, lp=20, m=60,
                                            density=0.03, seed=5),
    # one block of at least 128 rows: the library GEMM paths (F1 products,
    # inverse and eigenvalue set-up)
    "maxcut150": lambda: PR.maxcut(150, 0.1, 4),
    "dense_n130_m60": lambda: PR.dense_random(n=130, m=60, seed=6),
}


@pytest.mark.parametrize("name", list(SOLVE_CASES))
def test_full_solve_matches_oracle(name):
    d = SOLVE_CASES[name]()
    r = sa.Solver(0).load(d).solve()
    o = O.Problem(d).solve()
    print(name, r.phase, r.iterations, o.iterations, r.primal_obj, o.primal_obj)
    assert r.phase == o.phase
    assert abs(r.iterations - o.iterations) <= 1
    assert r.primal_obj == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8)
    assert r.dual_obj == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8)
    if o.phase == "pdOPT":
        assert r.relative_gap <= 1e-7
        assert r.primal_error <= 1e-7 and r.dual_error <= 1e-7


def test_init_point_solve(golden_dir):
    import os
    s = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    s.set_init(np.array([5.9, -1.375, -1.375, 1.0]), np.array([11.0, 0.0, 0.0, 9.0]),
               -np.array([0.0, -4.0, 0.0]))
    r = s.solve()
    assert r.phase == "pdOPT"
    assert r.primal_obj == pytest.approx(-41.9, rel=1e-6)


@pytest.mark.parametrize("per_block,dense", [(60, False), (50, True)])
def test_sparse_schur_decision_shift(per_block, dense):
    # block-separated constraints: m = 120 leaves the reference's dense path
    # (Chordal::ordering_bMat), m = 100 stays on it.  Off the dense path the
    # reference factors bMat + 1e-10 I (sdpa_chordal.cpp:526-534); on a
    # diagonal matrix that is a closed form: L_ii = sqrt(d_i + 1e-10), against
    # the dense rule's sqrt(d_i) and 1e50 for a zero pivot (d < 1e-14 -> 1e100)
    d = PR.block_separated(per_block=per_block)
    g = ctx_for(d)
    assert bool(g.info[9]) == dense == O.Problem(d).schur_dense
    m = d.m
    dg = np.ones(m)
    dg[5] = 0.0
    dg[50] = 1e-13
    g.set("B", np.diag(dg).reshape(-1, order="F"))
    assert g.factor_bmat() == 0
    L = np.diag(g.bmat())
    want = np.sqrt(dg + 1e-10) if not dense else np.where(dg < 1e-14, 1e50, np.sqrt(dg))
    np.testing.assert_allclose(L, want, rtol=1e-14)
    # the whole solve against the oracle (which applies the same rule)
    r = sa.Solver(0).load(d).solve()
    o = O.Problem(d).solve()
    print(d.name, r.phase, r.iterations, o.iterations, r.primal_obj, o.primal_obj)
    assert r.phase == o.phase == "pdOPT"
    assert abs(r.iterations - o.iterations) <= 1
    assert r.primal_obj == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8)
    assert r.dual_obj == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8)


@pytest.mark.parametrize("per_block,couple", [(64, 0), (64, 3), (40, 2)])
def test_sparse_branch_tile_factorization(per_block, couple):
    """The sparse-Schur branch factored by tiles of the ordered matrix: 20
    blocks of 64 local constraints (block-diagonal Schur pattern) or 30 of 40
    (tiles straddle blocks: fill, several levels), plus `couple` constraints
    touching every block (an arrow).  The loader leaves
    the dense path like the oracle (the aggregate-pattern rule) and takes the
    tile factorization (its flops far below the dense m^3 / 3).  The factor
    and solve of bMat + 1e-10 I at two iterates: dy against numpy's solve of
    the same shifted matrix (B from the device, its lower triangle), relative
    error <= 1e-9; an indefinite matrix of the same pattern fails the
    factorization; the whole solve against the oracle's (which factors the
    same shifted matrix densely)."""
    d = PR.block_separated(n_blocks=20 if per_block == 64 else 30, per_block=per_block,
                           couple=couple, seed=31)
    g = ctx_for(d)
    m = d.m
    assert int(g.info[9]) == 0 and not O.Problem(d).schur_dense
    print(d.name, "tile factorization", int(g.info[12]), "tiles", int(g.info[13]), "per side",
          int(g.info[14]), "levels", int(g.info[15]))
    assert int(g.info[12]) == 1
    assert int(g.info[13]) < int(g.info[14]) * (int(g.info[14]) + 1) // 2  # not all tiles
    rng = np.random.default_rng(couple)
    for lam in (10.0, 0.3):
        g.identity_start(lam)
        g.make_bmat()
        B = g.bmat()
        Bs = np.tril(B) + np.tril(B, -1).T
        x = rng.standard_normal(m)
        g.set("g", (Bs + 1e-10 * np.eye(m)) @ x)
        assert g.factor_bmat() == 0
        g.solve_bmat()
        dy = g.get("dy")
        want = np.linalg.solve(Bs + 1e-10 * np.eye(m), (Bs + 1e-10 * np.eye(m)) @ x)
        rel = np.linalg.norm(dy - want) / np.linalg.norm(want)
        print(f"lambda={lam}: solve relative error {rel:.2e}")
        assert rel <= 1e-9
        assert np.array_equal(g.bmat(), B)  # the assembled matrix is left as it was
    # an indefinite matrix with the same pattern
    Bn = Bs.copy()
    Bn[np.diag_indices(m)] *= -1.0
    g.set("B", Bn.reshape(-1, order="F"))
    assert g.factor_bmat() > 0
    g.close()
    r = sa.Solver(0).load(d).solve()
    o = O.Problem(d).solve()
    print(d.name, r.phase, r.iterations, o.iterations, r.primal_obj, o.primal_obj)
    assert r.phase == o.phase == "pdOPT"
    assert abs(r.iterations - o.iterations) <= 1
    assert r.primal_obj == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8)
    assert r.dual_obj == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8)
    assert r.relative_gap <= 1e-7
