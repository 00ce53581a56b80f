"""The batched solver's no-allocation entry point (sdpa_batch_solve_noalloc,
include/sdpa_batch.h) against the allocating one (sdpa_batch_solve) on the
same instances, through the C ABI: two calls in a row, the second with a
smaller batch, each compared with sdpa_batch_solve's results.  Both run the
same kernel on the same data; each instance's arithmetic does not depend on
where its work slab lives, so phases, iteration counts and every result
value agree exactly.  A batch larger than the fixed scratch array is
rejected on the host (-9) before anything runs on the device."""
import numpy as np
import pytest

from sdpa_amd import problems as PR

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sdpa_amd")
B = pytest.importorskip("sdpa_amd.batch")


def same(r, q):
    assert list(r.phase) == list(q.phase)
    assert (r.iterations == q.iterations).all()
    for f in ("primal_obj", "dual_obj", "primal_error", "dual_error", "mu", "failure"):
        np.testing.assert_array_equal(getattr(r, f), getattr(q, f), err_msg=f)


def test_noalloc_matches_allocating_entry_point():
    big = PR.maxcut_batch(600, nmin=8, nmax=64, p=0.2, seed=21)
    small = PR.maxcut_batch(40, nmin=20, nmax=40, p=0.3, seed=22)
    for insts in (big, small):
        want = B.solve(insts)
        got = B.solve_noalloc(insts)
        print(f"{len(insts)} instances: phases {sorted(set(got.phase))}, "
              f"iterations {got.iterations.min()}-{got.iterations.max()}", flush=True)
        same(got, want)
    # the per-instance lambda and parameters reach the kernel the same way
    lam = np.geomspace(1.0, 1e4, len(small))
    same(B.solve_noalloc(small, lambdas=lam, maxIteration=6),
         B.solve(small, lambdas=lam, maxIteration=6))


def test_noalloc_rejects_a_batch_beyond_its_scratch():
    insts = PR.maxcut_batch(4096)
    with pytest.raises(sa.GpuError, match=r"\(-9\)"):
        B.solve_noalloc(insts)
