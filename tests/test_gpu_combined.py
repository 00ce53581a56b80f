"""The combined entry points against the separate calls they replace, on the
same device state: sdpa_gpu_min_eig_dots = sdpa_gpu_min_eig + sdpa_gpu_dots
(StepLength::MehrotraCorrector, sdpa_parts.cpp:493-625) and
sdpa_gpu_update_end = sdpa_gpu_inv_chol + sdpa_gpu_dots + sdpa_gpu_residuals
(sdpa_solve.cpp:167-190).  The same kernels run in the same order, so the
results are identical to the last bit; an indefinite X makes both report the
failed factorization."""
import numpy as np
import pytest

import sdpa_amd as sa
from sdpa_amd import problems as PR
from tests.test_gpu_parity import packed_random

pytestmark = pytest.mark.gpu

CASES = {
    "maxcut60": lambda: PR.maxcut(60, 0.2, 3),
    # several SDP blocks and an LP block (the LP slots of both calls)
    "multiblock_lp": lambda: PR.multi_block(n_sdp_blocks=6, nmin=3, nmax=30, lp=12, m=40,
                                            density=0.05, seed=9),
}
STEP_CODES = ("CDX", "BDY", "XDZ", "DXZ", "DXDZ", "BY", "CX")
END_CODES = ("XZ", "CX", "BY", "TRZ", "TRX")


def _direction_state(d):
    """A context with X, Z, Zinv set and a predictor direction Dx, Dz."""
    g = sa.GpuContext(0).load(d)
    X, Z, Zi = packed_random(d, np.random.default_rng(4))
    g.set("X", X); g.set("Z", Z)
    assert g.inv_chol() == 0
    g.make_gvec(0, 0.1, 1)
    g.make_bmat()
    assert g.factor_bmat() == 0
    g.solve_bmat()
    g.make_direction(1)
    return g, X


@pytest.mark.parametrize("name", list(CASES))
def test_min_eig_dots_equals_separate_calls(name):
    g, _ = _direction_state(CASES[name]())
    lx, lz = g.min_eig()
    q = g.dots(*STEP_CODES)
    lx2, lz2, q2 = g.min_eig_dots(*STEP_CODES)
    assert (lx2, lz2) == (lx, lz)
    np.testing.assert_array_equal(q2, q)
    g.close()


@pytest.mark.parametrize("name", list(CASES))
def test_update_end_equals_separate_calls(name):
    d = CASES[name]()
    g, X = _direction_state(d)
    # half the step to the boundary of the cone (X + a Dx >= 0 for
    # a <= -1 / lambda_min(L^-1 Dx L^-T))
    lx, lz = g.min_eig()
    ap = 0.5 * min(1.0, -1.0 / lx) if lx < 0 else 0.5
    ad = 0.5 * min(1.0, -1.0 / lz) if lz < 0 else 0.5
    g.update(ap, ad)
    assert g.inv_chol() == 0
    icx, icz, zi = g.get("invCholX"), g.get("invCholZ"), g.get("Zinv")
    o = g.dots(*END_CODES)
    nP, nD = g.residuals()
    r, o2, nP2, nD2 = g.update_end(*END_CODES)
    assert r == 0
    np.testing.assert_array_equal(o2, o)
    assert (nP2, nD2) == (nP, nD)
    np.testing.assert_array_equal(g.get("invCholX"), icx)
    np.testing.assert_array_equal(g.get("invCholZ"), icz)
    np.testing.assert_array_equal(g.get("Zinv"), zi)
    # an indefinite X block: both report the failed factorization
    Xbad = g.get("X").copy()
    n0 = int(d.block_struct[0])
    Xbad[0] = -abs(Xbad[0]) - 10.0 * np.abs(Xbad[: n0 * n0]).max()
    g.set("X", Xbad)
    assert g.update_end(*END_CODES)[0] == 1
    assert g.inv_chol() == 1
    g.close()
