"""CPU checks of the lambda* model's oracle (oracle/mlp_oracle.py) and of the
host-side parts of sdpa_amd.lambda_mlp (features, initial point).

The oracle's gradient is pinned by central finite differences of its loss;
its Adam step by the update written out for one coordinate.  Features are
checked through properties the definitions fix (scaling C by 10 raises the
log10 norm of C by exactly 1; the LP flag and block sizes)."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from sdpa_amd import lambda_mlp as LM
from sdpa_amd import problems as PR


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    nin, nh, n = 5, 7, 40
    th = rng.standard_normal(MO.nparams(nin, nh)) * 0.5
    X = rng.standard_normal((n, nin))
    t = rng.standard_normal(n)
    loss, g = MO.loss_grad(th, X, t, nin, nh)
    h = 1e-6
    fd = np.zeros_like(th)
    for k in range(th.shape[0]):
        e = np.zeros_like(th)
        e[k] = h
        fd[k] = (MO.loss_grad(th + e, X, t, nin, nh)[0] - MO.loss_grad(th - e, X, t, nin, nh)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)
    # the loss is the mean squared error of the forward pass
    y, _, _ = MO.forward(th, X, nin, nh)
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-14)


def test_oracle_adam_step_one_coordinate():
    th, g = np.array([1.0]), np.array([0.5])
    m1, m2 = np.zeros(1), np.zeros(1)
    th1, m1, m2 = MO.adam_step(th, g, m1, m2, 1, 0.01)
    # first step: m1 = 0.05, m2 = 0.00025, bias-corrected to g and g^2
    assert m1[0] == pytest.approx(0.05) and m2[0] == pytest.approx(0.00025)
    assert th1[0] == pytest.approx(1.0 - 0.01 * 0.5 / (0.5 + 1e-8), rel=1e-12)


def test_block_features_properties():
    d = PR.multi_block(n_sdp_blocks=3, nmin=3, nmax=6, lp=5, m=12, density=0.2, seed=9)
    F = LM.block_features(d)
    assert F.shape == (d.n_block, LM.NIN) and np.isfinite(F).all()
    assert list(F[:, 1]) == [0.0, 0.0, 0.0, 1.0]  # LP flag
    assert np.allclose(np.exp(F[:, 0]), np.abs(d.block_struct))
    d10 = PR.SdpData(**{**d.__dict__, "v": np.where(d.k == 0, 10 * d.v, d.v)})
    F10 = LM.block_features(d10)
    np.testing.assert_allclose(F10[:, 3] - F[:, 3], 1.0, atol=1e-12)
    np.testing.assert_allclose(np.delete(F10, 3, axis=1), np.delete(F, 3, axis=1))


def test_init_point_layout():
    d = PR.multi_block(n_sdp_blocks=2, nmin=2, nmax=3, lp=4, m=6, density=0.3, seed=2)
    lam = np.array([2.0, 3.0, 5.0])
    X, Z, y = LM.init_point(d, lam)
    assert X.shape[0] == d.packed_size() and (X == Z).all() and (y == 0).all()
    o = 0
    for l, s in enumerate(d.block_struct):
        s = int(s)
        if s >= 2:
            np.testing.assert_array_equal(X[o:o + s * s].reshape(s, s), lam[l] * np.eye(s))
            o += s * s
        else:
            np.testing.assert_array_equal(X[o:o + abs(s)], lam[l])
            o += abs(s)


def test_scaled_family_is_seeded():
    a, b = LM.scaled_problem(1234), LM.scaled_problem(1234)
    assert a.name == b.name and np.array_equal(a.v, b.v) and np.array_equal(a.c, b.c)


def test_label_rule():
    P, F = "pdOPT", "pFEAS"
    runs = [(0.1, 1, F), (1.0, 20, P), (10.0, 9, P), (100.0, 9, P), (1e3, 9, P), (1e4, 12, P)]
    # 1.0 is next to a failure; 10, 100, 1000 tie: the middle one
    assert LM.label_from_runs(runs) == 100.0
    runs = [(0.1, 1, F), (1.0, 5, P), (10.0, 1, F)]
    assert LM.label_from_runs(runs) == 1.0  # no safe choice: any pdOPT
    assert LM.label_from_runs([(1.0, 3, F)]) is None
