"""One solver reloaded with many problems in turn: buffers are reallocated
at every load, the device keeps its copies of the factorization's, inverse's
and GEMMs' descriptor arrays across calls (and drops them all when there are
too many), and every solve must give what a fresh solver gives.  SDPA::solve
(sdpa_solve.cpp:59-233) on each problem."""
import numpy as np
import pytest

import sdpa_amd as sa
from sdpa_amd import problems as PR

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    for k in range(14):
        n = 30 + 7 * k
        out.append(PR.maxcut(n, 0.15, 100 + k))
    for k in range(6):
        out.append(PR.multi_block(n_sdp_blocks=4 + 2 * k, nmin=3, nmax=20 + 3 * k, lp=10 + k,
                                  m=40 + 10 * k, density=0.03, seed=200 + k))
    return out


def test_reloaded_solver_matches_fresh_solvers():
    probs = _problems()
    reused = sa.Solver(0)
    for d in probs:
        r = reused.load(d).solve()
        f = sa.Solver(0)
        rf = f.load(d).solve()
        f.close()
        print(d.name, r.phase, r.iterations, r.primal_obj, rf.primal_obj)
        assert r.phase == rf.phase, d.name
        assert r.iterations == rf.iterations, d.name
        np.testing.assert_allclose([r.primal_obj, r.dual_obj], [rf.primal_obj, rf.dual_obj],
                                   rtol=1e-12, err_msg=d.name)
    reused.close()
