"""Batched small-SDP solver (include/sdpa_batch.h): every instance of one
launch against the oracle's SDPA::solve on the same instance and lambda.

Tolerances (FP64 on both sides, different summation orders): the oracle's
phase; iteration count within one; objectives within 1e-6 relative (1e-8
absolute), except on the unbounded side of a pUNBD/dUNBD instance, where
only the bound crossing is fixed; at pdOPT the relative gap and both feasibility errors at the 1e-7
stopping level.  Instances cover the Jacobi (n <= 20) and Lanczos (n > 20)
step-length branches, trace rows (m = n + 1), dense C (Lovasz), ownership
ties, the reference's example1, a pUNBD and a dUNBD case, n and m at the
limits (64, 72), and per-instance lambda.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import problems as PR

from test_gpu_parity import sparse_mixed

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sdpa_amd")
B = pytest.importorskip("sdpa_amd.batch")


def unbounded():
    # F_1 = -I with c_1 = 1: the user dual asks tr(Y) = -1 (infeasible) and
    # the primal objective x_1 decreases without bound.  (Dependent
    # constraints, F_i = F_j, are left out: there the Schur pivot cancels to
    # rounding level and whether rATL_dpotrfL's adjustment fires -- bounded
    # step or not -- depends on the last bits of the iterate.)
    recs = [(0, 1, 1, 1, 1.0), (0, 1, 2, 2, 2.0), (0, 1, 3, 3, 3.0), (1, 1, 1, 1, -1.0),
            (1, 1, 2, 2, -1.0), (1, 1, 3, 3, -1.0), (2, 1, 1, 2, 1.0), (3, 1, 2, 3, 1.0)]
    K, L, I, J, V = PR._records(recs)
    return PR._make(3, [3], [1, 0, 0], K, L, I, J, V, "unbounded_n3_m3")


def scaled_maxcut():
    d = PR.maxcut(20, 0.3, 7)
    d.v = d.v * np.where(d.k == 0, 1e3, 1.0)
    return d


def at_limits():
    # n = 64 with 8 extra rows: m = 72
    d = PR.maxcut(64, 0.2, 9)
    rng = np.random.default_rng(9)
    K, I, J, V = [d.k], [d.i], [d.j], [d.v]
    for t in range(8):
        a = rng.choice(64, 6, replace=False) + 1
        K.append(np.full(6, 65 + t, np.int32)); I.append(a.astype(np.int32))
        J.append(a.astype(np.int32)); V.append(np.ones(6))
    c = np.concatenate([d.c, np.full(8, 6.0)])
    K = np.concatenate(K); I = np.concatenate(I); J = np.concatenate(J); V = np.concatenate(V)
    return PR._make(72, [64], c, K, np.ones_like(K), I, J, V, "maxcut_n64_m72")


def check(d, o, r, t):
    msg = f"{d.name} n={d.block_struct[0]} m={d.m}"
    assert r.phase[t] == o.phase, msg
    assert abs(int(r.iterations[t]) - o.iterations) <= 1, msg
    if o.phase in ("pUNBD", "dUNBD"):
        # a diverging iterate: the objective on the unbounded side is only
        # fixed to be past the bound (1e5) at the iteration that detects it
        if o.phase == "pUNBD":
            assert r.primal_obj[t] <= -1e5 and o.primal_obj <= -1e5, msg
        else:
            assert r.dual_obj[t] >= 1e5 and o.dual_obj >= 1e5, msg
        return
    assert r.primal_obj[t] == pytest.approx(o.primal_obj, rel=1e-6, abs=1e-8), msg
    assert r.dual_obj[t] == pytest.approx(o.dual_obj, rel=1e-6, abs=1e-8), msg
    if o.phase == "pdOPT":
        assert r.relative_gap[t] <= 1e-7, msg
        assert r.primal_error[t] <= 1e-7 and r.dual_error[t] <= 1e-7, msg
        assert r.failure[t] == 0, msg


def instances():
    insts = PR.maxcut_batch(16, nmin=8, nmax=64, p=0.2, seed=4)
    insts += [PR.lovasz(12, 0.3, 5), PR.lovasz(16, 0.25, 6), PR.example1(),
              sparse_mixed(30, 60, 20, 5), unbounded(), scaled_maxcut(), at_limits()]
    return insts


@pytest.mark.parametrize("threads", [256, 128])
def test_batch_matches_oracle(threads):
    insts = instances()
    print(f"batch of {len(insts)} instances, {threads} threads", flush=True)
    r = B.solve(insts, threads=threads)
    for t, d in enumerate(insts):
        o = O.Problem(d).solve()
        print(d.name, r.phase[t], o.phase, r.iterations[t], o.iterations, r.primal_obj[t],
              o.primal_obj)
        check(d, o, r, t)
    assert r.kernel_ms > 0


def test_batch_per_instance_lambda():
    insts = [scaled_maxcut(), PR.maxcut(30, 0.2, 3), PR.lovasz(14, 0.3, 8)]
    lams = np.array([1e4, 10.0, 1e3])
    r = B.solve(insts, lambdas=lams)
    for t, d in enumerate(insts):
        o = O.Problem(d).solve(O.default_param(lambdaStar=float(lams[t])))
        check(d, o, r, t)


def test_batch_parameters_and_repeat():
    # maxIteration caps the loop (noINFO-or-feasible phase, iterations = cap);
    # a larger batch (instances repeated, more than one per work-group) gives
    # every copy the same answer
    insts = PR.maxcut_batch(6, nmin=10, nmax=40, seed=12)
    r = B.solve(insts, maxIteration=4)
    for t, d in enumerate(insts):
        o = O.Problem(d).solve(O.default_param(maxIteration=4))
        assert r.iterations[t] == o.iterations == 4
        assert r.phase[t] == o.phase
        assert r.primal_obj[t] == pytest.approx(o.primal_obj, rel=1e-9)
    many = insts * 100  # 600 instances
    rm = B.solve(many)
    r1 = B.solve(insts)
    for t in range(len(many)):
        assert rm.iterations[t] == r1.iterations[t % 6]
        assert rm.primal_obj[t] == r1.primal_obj[t % 6]
        assert rm.dual_obj[t] == r1.dual_obj[t % 6]


def test_batch4096_matches_oracle_fixture():
    """All 4096 instances of the batched workload (problems.maxcut_batch(4096))
    against the oracle's solves committed in tests/golden/oracle_batch4096.json
    (made by make_oracle_fixtures.py).  Where both end in the same phase:
    iterations within one, objectives within 1e-6 relative (1e-8 absolute).
    The phase may differ only between pdOPT and pdFEAS (an instance whose
    Schur matrix turns numerically singular near the end, where the pivot
    rule's 1e-14 threshold decides on the last bits), on at most 4 of 4096."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "oracle_batch4096.json")))
    insts = PR.maxcut_batch(4096)
    r = B.solve(insts)
    differ = []
    for t in range(len(insts)):
        if r.phase[t] != gold["phase"][t]:
            differ.append((t, r.phase[t], gold["phase"][t]))
            continue
        assert abs(int(r.iterations[t]) - gold["iterations"][t]) <= 1, t
        assert r.primal_obj[t] == pytest.approx(gold["primal_obj"][t], rel=1e-6, abs=1e-8), t
        assert r.dual_obj[t] == pytest.approx(gold["dual_obj"][t], rel=1e-6, abs=1e-8), t
        if r.phase[t] == "pdOPT":
            assert r.relative_gap[t] <= 1e-7, t
    print(f"batch4096: {len(insts) - len(differ)} of {len(insts)} in the oracle's phase; "
          f"differ: {differ}")
    assert len(differ) <= 4
    assert all({a, b} == {"pdOPT", "pdFEAS"} for _, a, b in differ)


def test_batch_mixed_precision_matches_oracle():
    """precision 1 (FP32 Schur factor of the equilibrated matrix, FP64
    refinement to dsposv's test): the oracle's FP64 results to the same
    tolerances as the FP64 launch, with the FP32 factorization taken on most
    iterations."""
    insts = instances()
    r = B.solve_cfg(insts, cfg=B.config(precision=1))
    for t, d in enumerate(insts):
        o = O.Problem(d).solve()
        print(d.name, r.phase[t], o.phase, r.iterations[t], o.iterations, r.primal_obj[t],
              o.primal_obj, r.stats[t])
        check(d, o, r, t)
    st = r.stats
    assert st[:, 0].sum() > 0.5 * (st[:, 0] + st[:, 1]).sum()  # mostly FP32 factorizations
    assert (st[:, 2] >= st[:, 0]).all()  # at least one refinement step per FP32 solve pair


def test_batch_mu_switch_and_fp64_identity():
    """precision 1 with a switch above every mu never factors in FP32 and
    gives the FP64 launch's results bit for bit; a switch inside the run
    factors in FP32 first and in FP64 after."""
    insts = PR.maxcut_batch(24, nmin=8, nmax=48, seed=21)
    r64 = B.solve_cfg(insts, cfg=B.config(precision=0))
    rhi = B.solve_cfg(insts, cfg=B.config(precision=1, mu_switch=1e30))
    assert (rhi.stats[:, 0] == 0).all()
    assert np.array_equal(rhi.primal_obj, r64.primal_obj)
    assert np.array_equal(rhi.iterations, r64.iterations)
    rmid = B.solve_cfg(insts, cfg=B.config(precision=1, mu_switch=1e-2))
    assert (rmid.stats[:, 0] > 0).all() and (rmid.stats[:, 1] > 0).all()
    for t, d in enumerate(insts):
        o = O.Problem(d).solve()
        check(d, o, rmid, t)


def test_batch_launch_configurations():
    """Every heavy-pair threshold and queue order of the autotuner's grid
    solves to the oracle's results; the queue order does not change any
    instance's arithmetic (bitwise equal)."""
    insts = instances()
    ref = [O.Problem(d).solve() for d in insts]
    for heavy in (16, 256):
        r0 = B.solve_cfg(insts, cfg=B.config(threads=128, heavy=heavy, order=0))
        r1 = B.solve_cfg(insts, cfg=B.config(threads=128, heavy=heavy, order=1))
        assert np.array_equal(r0.primal_obj, r1.primal_obj)
        assert np.array_equal(r0.iterations, r1.iterations)
        for t, d in enumerate(insts):
            check(d, ref[t], r0, t)


def test_batch_autotune():
    """The autotuner times the 20 candidates of its grid on a sample and
    returns the fastest; the batch solved under it matches the oracle
    fixture's first 512 instances."""
    import json
    import os
    insts = PR.maxcut_batch(4096)[:512]
    best, table = B.autotune(insts, sample=128)
    print("autotune best", best.as_dict())
    print(table)
    assert table.shape == (20, 4) and (table[:, 3] > 0).all()
    assert best.threads in (128, 256) and best.heavy in (16, 32, 64, 128, 256)
    assert table[:, 3].min() == pytest.approx(
        table[(table[:, 0] == best.threads) & (table[:, 1] == best.heavy) &
              (table[:, 2] == best.order), 3][0])
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "oracle_batch4096.json")))
    r = B.solve_cfg(insts, cfg=best)
    same = [t for t in range(len(insts)) if r.phase[t] == gold["phase"][t]]
    assert len(same) >= len(insts) - 1
    for t in same:
        assert abs(int(r.iterations[t]) - gold["iterations"][t]) <= 1, t
        assert r.primal_obj[t] == pytest.approx(gold["primal_obj"][t], rel=1e-6, abs=1e-8), t


def test_batch4096_mixed_precision_matches_oracle_fixture():
    """The 4096-instance workload with the FP32 Schur factorization and FP64
    refinement against the oracle's FP64 solves (same criteria as the FP64
    launch's fixture test)."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "oracle_batch4096.json")))
    insts = PR.maxcut_batch(4096)
    r = B.solve_cfg(insts, cfg=B.config(precision=1))
    differ = []
    for t in range(len(insts)):
        if r.phase[t] != gold["phase"][t]:
            differ.append((t, r.phase[t], gold["phase"][t]))
            continue
        assert abs(int(r.iterations[t]) - gold["iterations"][t]) <= 1, t
        assert r.primal_obj[t] == pytest.approx(gold["primal_obj"][t], rel=1e-6, abs=1e-8), t
        assert r.dual_obj[t] == pytest.approx(gold["dual_obj"][t], rel=1e-6, abs=1e-8), t
        if r.phase[t] == "pdOPT":
            assert r.relative_gap[t] <= 1e-7, t
    st = r.stats
    print(f"batch4096 mixed: {len(insts) - len(differ)} in the oracle's phase; differ {differ}; "
          f"FP32 factorizations {st[:, 0].sum():.0f}, FP64 {st[:, 1].sum():.0f}, "
          f"refinement steps {st[:, 2].sum():.0f}, stalls {st[:, 3].sum():.0f}; "
          f"kernel {r.kernel_ms:.1f} ms")
    assert len(differ) <= 4
    assert all({a, b} == {"pdOPT", "pdFEAS"} for _, a, b in differ)
