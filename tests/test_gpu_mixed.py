"""Mixed-precision Schur factorization (precision 1 of sdpa_gpu_factor_bmat:
FP32 factor of the column-equilibrated Schur matrix on FP32 MFMA, FP64
iterative refinement of every solve) through the C ABI.

Tolerances: a refined solution must be as accurate as the FP64 path's
(1e-9 relative to the solution norm, against numpy's FP64 solve, on
well-conditioned matrices; on an ill-conditioned one the protocol -- stall,
FP64 refactorization, solve -- must end in a backward-stable solution,
||A x - b|| <= 1e-12 ||A|| ||x||).  Whole solves: the oracle's phase,
iteration count within one, objectives within 1e-6 relative, gap and
feasibility at the 1e-7 stopping level.  The FP64 Schur matrix must be left
untouched by the mixed factorization (the refinement residuals use it).
"""
import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import problems as PR

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sdpa_amd")


def spd(n, rng, cond):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1, cond, n)) @ q.T


def stats(g):
    return dict(zip(("factors", "fallbacks", "solves", "steps", "stalls", "current"),
                    (int(v) for v in g.mixed_stats())))


@pytest.mark.parametrize("m,cond", [(700, 1e3), (1100, 1e5)])
def test_mixed_factor_and_refined_solve(m, cond):
    rng = np.random.default_rng(m)
    d = PR.maxcut(m, 0.5, seed=3)  # any problem with this m; only bMat is used
    g = sa.GpuContext(0).load(d)
    # a badly scaled SPD matrix: the equilibration must undo the scaling
    s = np.geomspace(1e-3, 1e3, m)
    A = spd(m, rng, cond) * np.outer(s, s)
    g.set("B", A.reshape(-1, order="F"))
    assert g.factor_bmat(1) == 0
    st = stats(g)
    assert st["current"] == 1 and st["factors"] == 1 and st["fallbacks"] == 0
    np.testing.assert_array_equal(g.bmat(), A)  # the FP64 matrix is intact
    # reference: numpy on the equilibrated matrix (well conditioned), scaled back;
    # errors are measured in the equilibrated variables S^-1 x
    d_ = 1.0 / np.sqrt(np.diag(A))
    Ae = A * np.outer(d_, d_)
    for rep in range(2):  # two right-hand sides with one factorization
        rhs = rng.standard_normal(m) / s
        g.set("g", rhs)
        assert g.solve_bmat() == 0
        x = g.get("dy")
        want = d_ * np.linalg.solve(Ae, d_ * rhs)
        assert np.linalg.norm((x - want) / d_) <= 1e-9 * np.linalg.norm(want / d_)
    st2 = stats(g)
    assert st2["solves"] == 2 and st2["stalls"] == 0
    # an FP32 factor alone is not FP64-accurate: every solve refines at least once
    assert st2["steps"] >= 4


def test_mixed_ill_conditioned_falls_back_to_fp64():
    rng = np.random.default_rng(11)
    m = 800
    d = PR.maxcut(m, 0.5, seed=3)
    g = sa.GpuContext(0).load(d)
    A = spd(m, rng, 1e12)
    g.set("B", A.reshape(-1, order="F"))
    rhs = rng.standard_normal(m)
    assert g.factor_bmat(1) == 0
    g.set("g", rhs)
    rc = g.solve_bmat()
    st = stats(g)
    if st["current"] == 1:
        # FP32 cannot resolve cond 1e12: the refinement must report a stall
        assert rc == 2 and st["stalls"] == 1
        assert g.factor_bmat(0) == 0
        rc = g.solve_bmat()
    else:
        assert st["fallbacks"] == 1  # the FP32 factorization itself failed
    assert rc == 0
    x = g.get("dy")
    res = np.linalg.norm(A @ x - rhs)
    assert res <= 1e-12 * np.linalg.norm(A, 2) * np.linalg.norm(x)


def test_mixed_small_m_uses_fp64():
    d = PR.maxcut(100, 0.1, 1)
    g = sa.GpuContext(0).load(d)
    A = spd(100, np.random.default_rng(1), 10.0)
    g.set("B", A.reshape(-1, order="F"))
    assert g.factor_bmat(1) == 0
    st = stats(g)
    assert st["current"] == 0 and st["fallbacks"] == 1


@pytest.mark.parametrize("n", [60, 70])
def test_mixed_full_solve_matches_oracle(n):
    d = PR.lovasz(n, 0.3, 2)  # m = 550, 753: inside the mixed path's range
    s = sa.Solver(0).load(d).set_param(precision=1)
    r = s.solve()
    st = s.mixed_stats()
    o = O.Problem(d).solve()
    print(n, d.m, r.phase, r.iterations, o.iterations, r.primal_obj, o.primal_obj, st)
    assert st["factors"] >= 1  # the mixed path was taken
    assert r.phase == o.phase == "pdOPT"
    assert abs(r.iterations - o.iterations) <= 1
    assert r.primal_obj == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8)
    assert r.dual_obj == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8)
    assert r.relative_gap <= 1e-7
    assert r.primal_error <= 1e-7 and r.dual_error <= 1e-7


def test_mixed_beyond_float_solve_range_uses_fp64():
    # m = 13464 (Lovasz n = 300): more than 48 blocks of 256 rows, beyond the
    # pipelined float solves; the mixed factorization falls back to FP64 (and
    # the run stays there), so the run matches the FP64 run (two iterations)
    d = PR.lovasz(300, 0.3, 2)
    s = sa.Solver(0).load(d).prepare()
    r0 = s.set_param(precision=0, maxIteration=2).run()
    r1 = s.set_param(precision=1, maxIteration=2).run()
    st = s.mixed_stats()
    print(d.m, st, r0.primal_obj, r1.primal_obj)
    assert d.m == 13464
    assert r1.iterations == r0.iterations == 2
    assert st["factors"] == 0 and st["fallbacks"] >= 1
    assert r1.primal_obj == r0.primal_obj and r1.dual_obj == r0.dual_obj


def test_mixed_mu_switch_hands_over_to_fp64():
    # muSwitch > 0: the Schur matrix is factored in FP32 (with refinement)
    # while mu > muSwitch and in FP64 from the first iteration with mu at or
    # below it (SDPA::setPrecision).  Lovasz n = 60 (m = 550) starts at
    # mu = 1e4: with muSwitch = 1 about half of the 13 factorizations are
    # mixed; the result is the oracle's, as with FP64 throughout.
    d = PR.lovasz(60, 0.3, 2)
    o = O.Problem(d).solve()
    s = sa.Solver(0).load(d).set_param(precision=1, muSwitch=1.0)
    r = s.solve()
    st = s.mixed_stats()
    mus = [t[0] for t in r.trace]
    above = sum(1 for mu in mus[:r.iterations] if mu > 1.0)
    print(d.m, r.phase, r.iterations, st, "iterations with mu > 1:", above)
    # one mixed factorization (or FP64 fallback inside it) per iteration above the switch
    assert st["factors"] + st["fallbacks"] == above
    assert 0 < above < r.iterations
    assert st["current"] == 0  # the last factorization was FP64
    assert r.phase == o.phase == "pdOPT"
    assert abs(r.iterations - o.iterations) <= 1
    assert r.primal_obj == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8)
    assert r.dual_obj == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8)
    assert r.relative_gap <= 1e-7
    assert r.primal_error <= 1e-7 and r.dual_error <= 1e-7


def _sparse_f1_block():
    # one 30 x 30 block, constraints with 1..60 entries: F1 owners with many
    # entries and sparse partners (the FP32 path's densify and gather kernels)
    from test_gpu_parity import sparse_mixed
    return sparse_mixed()


FP32_CASES = {
    "dense_n70_m40": lambda: PR.dense_random(n=70, m=40, seed=6),
    "dense_n130_m200": lambda: PR.dense_random(n=130, m=200, seed=7),
    "sparse_f1_owners": _sparse_f1_block,
}


@pytest.mark.parametrize("name", list(FP32_CASES))
def test_fp32_assembly_matches_oracle(name):
    # precision 1 of sdpa_gpu_make_bmat: the F1/F2 products in FP32 (FP32
    # accumulation; the slices of the long A_i . G_j sums added in FP64), the
    # F3 terms in FP64.  Against the oracle's FP64 bMat: max |B32 - B64| at
    # most 1e-5 max |B64| -- X, Z^-1 and A_i rounded to FP32 (2^-24 relative)
    # and sums of up to n^2 terms; 1e-6 is typical.
    from test_gpu_parity import packed_random
    d = FP32_CASES[name]()
    rng = np.random.default_rng(1)
    X, Z, Zi = packed_random(d, rng)
    g = sa.GpuContext(0).load(d)
    g.set("X", X); g.set("Zinv", Zi)
    g.make_bmat(1)
    B32 = g.bmat()
    g.make_bmat(0)
    B64 = g.bmat()
    g.close()
    want = O.Problem(d).make_bmat(X, Zi)
    tril = np.tril_indices(d.m)
    scale = np.abs(want[tril]).max()
    err32 = np.abs(B32[tril] - want[tril]).max() / scale
    err64 = np.abs(B64[tril] - want[tril]).max() / scale
    print(f"{name}: FP32 assembly error {err32:.2e}, FP64 {err64:.2e} (relative to max |B|)")
    assert err32 <= 1e-5
    assert err64 <= 1e-12
    assert err32 > 0.0  # the FP32 products were used


def test_split_precision_dense_solve():
    # the split-precision path on an F1-dominated problem with m > 512 (the
    # mixed factorization's range): F1/F2 products in FP32, the Schur matrix
    # factored in FP32 with FP64 refinement until the first stall, FP64 after
    # it.  The answer is the FP64 path's (itself checked against the oracle):
    # pdOPT, objectives within 1e-6 relative, gap and feasibility at the 1e-7
    # stopping level.
    d = PR.dense_random(n=60, m=600, seed=8)
    s = sa.Solver(0).load(d).prepare()
    r0 = s.set_param(precision=0).run()
    r1 = s.set_param(precision=1).run()
    st = s.mixed_stats()
    print(d.m, r0.iterations, r1.iterations, r0.primal_obj, r1.primal_obj, st)
    assert st["factors"] >= 1
    assert r0.phase == r1.phase == "pdOPT"
    assert abs(r1.iterations - r0.iterations) <= 3
    assert r1.primal_obj == pytest.approx(r0.primal_obj, rel=1e-6, abs=1e-8)
    assert r1.dual_obj == pytest.approx(r0.dual_obj, rel=1e-6, abs=1e-8)
    assert r1.relative_gap <= 1e-7
    assert r1.primal_error <= 1e-7 and r1.dual_error <= 1e-7
