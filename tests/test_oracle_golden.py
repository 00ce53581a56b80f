"""The CPU oracle against the reference's own published known answers.

The SDPA manual prints the full iteration log and the final summary of
example1 (LP64/example1.dat-s) and example2 (LP64/example2.dat); both files
and the numbers are kept under tests/golden/.  Values below ~1e-10 in the
log are rounding noise (residual norms of feasible iterates) and step
lengths of 90 (= 0.9 * 100, a full dual step taken when lambda_min of the
scaled direction is above -1e-2) are decided by rounding-level directions;
those entries are not compared.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from sdpa_amd import dats


def _load(golden_dir, name):
    g = json.load(open(os.path.join(golden_dir, "manual_examples.json")))[name]
    d = dats.read(os.path.join(golden_dir, g["file"]))
    return g, d


def _close_printed(mine, ref, digits=2):
    if abs(ref) < 1e-10 and abs(mine) < 1e-10:
        return True
    return abs(mine - ref) <= 0.51 * 10 ** (np.floor(np.log10(abs(ref))) - (digits - 1)) + 1e-300


@pytest.mark.parametrize("name", ["example1", "example2"])
def test_manual_trace_and_summary(golden_dir, name):
    g, d = _load(golden_dir, name)
    out = O.Problem(d).solve()
    assert out.phase == g["phase"]
    assert out.iterations == g["iterations"]
    assert abs(out.primal_obj - g["objValPrimal"]) <= 1e-9 * abs(g["objValPrimal"])
    assert abs(out.dual_obj - g["objValDual"]) <= 1e-9 * abs(g["objValDual"])
    assert abs(out.relative_gap - g["relative_gap"]) <= 1e-6 * g["relative_gap"]
    assert abs(out.mu - g["mu"]) <= 1e-6 * g["mu"]
    assert out.primal_error < 1e-12 and out.dual_error < 1e-12
    mine = out.user_trace()
    assert len(mine) == len(g["trace"])
    for row_m, row_r in zip(mine, g["trace"]):
        it = int(row_r[0])
        ref = row_r[1:]
        # mu, thetaP, thetaD (1 decimal), objP, objD (2 decimals)
        for col, dg in ((0, 2), (1, 2), (2, 2), (3, 3), (4, 3)):
            if abs(ref[col]) < 1e-10:
                assert abs(row_m[col]) < 1e-10, (it, col)
                continue
            assert _close_printed(row_m[col], ref[col], dg), (name, it, col, row_m[col], ref[col])
        for col in (5, 6):
            if ref[col] > 50 or row_m[col] > 50:
                continue
            assert _close_printed(row_m[col], ref[col], 2), (name, it, col, row_m[col], ref[col])
        assert _close_printed(row_m[7], ref[7], 3)
    if "xVec" in g:
        np.testing.assert_allclose(out.x, g["xVec"], rtol=6e-4)


def test_dense_and_sparse_example1_agree(golden_dir):
    a = O.Problem(dats.read(os.path.join(golden_dir, "example1.dat"))).solve()
    b = O.Problem(dats.read(os.path.join(golden_dir, "example1.dat-s"))).solve()
    assert a.iterations == b.iterations
    assert a.primal_obj == pytest.approx(b.primal_obj, rel=1e-14)


def test_init_point_example1(golden_dir):
    """example1.ini-s (manual section 7): x0 = (0,-4,0), X0 = diag(11, 9),
    Y0 = [[5.9, -1.375], [-1.375, 1]].  Internal: Z0 = X0, X0_int = Y0,
    y0 = -x0 (sdpa_call.cpp:351, 385, 426)."""
    d = dats.read(os.path.join(golden_dir, "example1.dat-s"))
    P = O.Problem(d)
    Xint = np.array([5.9, -1.375, -1.375, 1.0])
    Zint = np.array([11.0, 0.0, 0.0, 9.0])
    y0 = -np.array([0.0, -4.0, 0.0])
    out = P.solve(X0=Xint, Z0=Zint, y0=y0)
    assert out.phase == "pdOPT"
    assert out.primal_obj == pytest.approx(-41.9, rel=1e-6)
    assert out.dual_obj == pytest.approx(-41.9, rel=1e-6)
