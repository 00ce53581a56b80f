"""Learned initial-point scale lambda* per block.

SDPA starts from X0 = Z0 = lambda* I, y0 = 0 (Solutions::initialize,
sdpa_dataset.cpp:47-74) with one lambda* from param.sdpa (default 100); a
per-block value can only enter through the initial point
(sdpa_solve.cpp:61-72).  This module predicts log10 lambda* for every block
of a problem from block features with a small MLP (8 features -> 64 ReLU ->
64 ReLU -> 1) that is trained and evaluated in this library's own HIP
kernels (include/sdpa_mlp.h).  The training labels come from solves the
solver itself runs: for each seeded training problem the lambda* of a grid
that reaches pdOPT in the fewest iterations.

The prediction only chooses the starting point: the stopping rules are the
solver's own, and `solve` falls back to the default lambda* (no initial
point) whenever the predicted start does not end in pdOPT.

    python -m sdpa_amd.lambda_mlp train --out sdpa_amd/data/lambda_mlp.npz \
        --report sdpa_amd/data/lambda_mlp_report.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import os
import time

import numpy as np

from . import Solver, lib
from . import problems as PR

NIN, NH = 8, 64
# the lambda* grid of the labels: 10^-2 .. 10^7 in half decades
LAMBDA_GRID = 10.0 ** (np.arange(-4, 15) / 2.0)
DEFAULT_LAMBDA = 100.0
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "lambda_mlp.npz")


def _lg(x):
    return float(np.log10(max(float(x), 1e-30)))


def block_features(d: PR.SdpData) -> np.ndarray:
    """One row per block: log n_b, LP flag, log m, log10 ||C_b||_F,
    log10 ||(A_1,b .. A_m,b)||_F, log10 ||b||_2, log10 max |b_i|,
    log(1 + constraints with data in the block).  Norms count every
    off-diagonal entry of an SDP block twice (both triangles)."""
    nb = d.n_block
    lp = d.block_struct < 0
    w = np.where((d.i != d.j) & ~lp[d.l - 1], 2.0, 1.0)
    sq = w * d.v * d.v
    c2 = np.zeros(nb)
    a2 = np.zeros(nb)
    np.add.at(c2, d.l[d.k == 0] - 1, sq[d.k == 0])
    np.add.at(a2, d.l[d.k > 0] - 1, sq[d.k > 0])
    # distinct constraints with data in each block: one pass over the (block,
    # constraint) keys
    sel = d.k > 0
    keys = np.unique((d.l[sel].astype(np.int64) - 1) * (d.m + 1) + d.k[sel])
    ncon = np.bincount(keys // (d.m + 1), minlength=nb).astype(float)
    b2 = float(np.linalg.norm(d.c))
    binf = float(np.max(np.abs(d.c))) if d.m else 0.0
    F = np.zeros((nb, NIN))
    for l in range(nb):
        n = abs(int(d.block_struct[l]))
        F[l] = [np.log(n), 1.0 if lp[l] else 0.0, np.log(d.m), 0.5 * _lg(c2[l]), 0.5 * _lg(a2[l]),
                _lg(b2), _lg(binf), np.log1p(ncon[l])]
    return F


def scaled_problem(seed: int) -> PR.SdpData:
    """Training and evaluation family: small MaxCut and Lovasz-theta
    instances and small multi-block problems (the generator of the
    200-block configuration at 3-20 blocks of 5-50, an LP block of 0-200,
    m = 50-400), whose cost matrix and right-hand side are scaled by
    independent factors 10^U(-3, 4), so the best lambda* varies over
    decades (and the default 100 can end in a wrong infeasibility or
    unboundedness verdict)."""
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(4))
    if kind == 3:
        d = PR.multi_block(n_sdp_blocks=int(rng.integers(3, 21)), nmin=5, nmax=50,
                           lp=int(rng.integers(0, 201)), m=int(rng.integers(50, 401)),
                           density=float(10.0 ** rng.uniform(-3, -1.5)),
                           seed=int(rng.integers(1 << 30)))
    elif kind == 0:
        d = PR.maxcut(int(rng.integers(10, 41)), float(rng.uniform(0.1, 0.5)), int(rng.integers(1 << 30)))
    elif kind == 1:
        d = PR.lovasz(int(rng.integers(8, 21)), float(rng.uniform(0.3, 0.5)), int(rng.integers(1 << 30)))
    else:
        d = PR.maxcut(int(rng.integers(10, 41)), float(rng.uniform(0.1, 0.5)), int(rng.integers(1 << 30)),
                      trace_row=True)
    sc, sb = 10.0 ** rng.uniform(-3, 4), 10.0 ** rng.uniform(-3, 4)
    v = d.v.copy()
    v[d.k == 0] *= sc
    return dataclasses.replace(d, v=v, c=d.c * sb, name=f"{d.name}_sc{sc:.3g}_sb{sb:.3g}")


def large_problem(seed: int) -> PR.SdpData:
    """Training family at the size of the 200-block configuration: 50-200
    SDP blocks of 5-50, an LP block of 200-1000, m = 1000-5000 at density
    1e-3; half of them with C and b scaled by 10^U(-2, 2)."""
    rng = np.random.default_rng(seed)
    d = PR.multi_block(n_sdp_blocks=int(rng.integers(50, 201)), nmin=5, nmax=50,
                       lp=int(rng.integers(200, 1001)), m=int(rng.integers(1000, 5001)),
                       density=1e-3, seed=int(rng.integers(1 << 30)))
    if rng.random() < 0.5:
        return d
    sc, sb = 10.0 ** rng.uniform(-2, 2), 10.0 ** rng.uniform(-2, 2)
    v = d.v.copy()
    v[d.k == 0] *= sc
    return dataclasses.replace(d, v=v, c=d.c * sb, name=f"{d.name}_sc{sc:.3g}_sb{sb:.3g}")


FAMILIES = {"small": scaled_problem, "large": large_problem}


def init_point(d: PR.SdpData, lam) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Packed X0 = Z0 = lambda_b I per block (internal convention), y0 = 0."""
    parts = []
    for l, s in enumerate(d.block_struct):
        s = int(s)
        if s >= 2:
            parts.append((lam[l] * np.eye(s)).reshape(-1, order="F"))
        else:
            parts.append(np.full(abs(s), lam[l]))
    X = np.concatenate(parts)
    return X, X.copy(), np.zeros(d.m)


class LambdaModel:
    """The MLP on the device (sdpa_mlp_* kernels) plus the feature
    normalization; parameters and normalization are saved together."""

    def __init__(self, device: int = 0, seed: int = 0):
        h = C.c_void_p()
        rc = lib().sdpa_mlp_create(device, NIN, NH, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"sdpa_mlp_create failed ({rc})")
        self.h = h
        self.np = int(lib().sdpa_mlp_nparams(h))
        rng = np.random.default_rng(seed)
        # He initialization of the weights, zero biases
        th = np.zeros(self.np)
        o = 0
        for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, 1)):
            th[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in)
            o += fan_in * fan_out + fan_out
        self.mu = np.zeros(NIN)
        self.sd = np.ones(NIN)
        self.set_params(th)

    def close(self):
        if getattr(self, "h", None):
            lib().sdpa_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def set_params(self, th):
        th = np.ascontiguousarray(th, np.float64)
        self._chk(lib().sdpa_mlp_set_params(self.h, th, th.shape[0]), "sdpa_mlp_set_params")

    def params(self):
        out = np.zeros(self.np)
        self._chk(lib().sdpa_mlp_get_params(self.h, out, self.np), "sdpa_mlp_get_params")
        return out

    def set_data(self, Xn, t):
        Xn = np.ascontiguousarray(Xn, np.float64)
        t = np.ascontiguousarray(t, np.float64)
        self._chk(lib().sdpa_mlp_set_data(self.h, Xn, t, Xn.shape[0]), "sdpa_mlp_set_data")

    def loss_grad(self):
        loss = C.c_double()
        g = np.zeros(self.np)
        self._chk(lib().sdpa_mlp_loss_grad(self.h, C.byref(loss), g.ctypes.data_as(C.c_void_p)),
                  "sdpa_mlp_loss_grad")
        return loss.value, g

    def train_steps(self, steps, lr):
        loss = C.c_double()
        self._chk(lib().sdpa_mlp_train(self.h, int(steps), float(lr), C.byref(loss)), "sdpa_mlp_train")
        return loss.value

    def predict_raw(self, Xn):
        Xn = np.ascontiguousarray(Xn, np.float64)
        y = np.zeros(Xn.shape[0])
        self._chk(lib().sdpa_mlp_predict(self.h, Xn, Xn.shape[0], y), "sdpa_mlp_predict")
        return y

    def fit(self, F, t, steps=3000, lr=3e-3):
        """Full-batch Adam on (features, log10 lambda*) in the device kernels."""
        self.mu = F.mean(0)
        self.sd = F.std(0)
        self.sd[self.sd < 1e-12] = 1.0
        self.set_data((F - self.mu) / self.sd, t)
        losses = []
        for _ in range(max(1, steps // 100)):
            losses.append(self.train_steps(100, lr))
        return losses

    def predict_log_lambda(self, F):
        return self.predict_raw((F - self.mu) / self.sd)

    def save(self, path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, theta=self.params(), mu=self.mu, sd=self.sd, nin=NIN, nh=NH)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device=0):
        z = np.load(path, allow_pickle=False)
        if int(z["nin"]) != NIN or int(z["nh"]) != NH:
            raise ValueError("weights do not match this model's shape")
        m = cls(device)
        m.set_params(z["theta"])
        m.mu, m.sd = z["mu"], z["sd"]
        return m


def label_from_runs(runs):
    """The label of one problem from its grid of solves [(lambda, iterations,
    phase)]: the fewest iterations among the lambda* that reach pdOPT and
    whose grid neighbours reach pdOPT too (a start next to a wrong verdict
    is a poor target), and of those the middle one; None if nothing
    reaches pdOPT."""
    ok = [ph == "pdOPT" for _, _, ph in runs]
    safe = [k for k in range(len(runs)) if ok[k] and (k == 0 or ok[k - 1]) and
            (k == len(runs) - 1 or ok[k + 1])]
    cand = safe or [k for k in range(len(runs)) if ok[k]]
    if not cand:
        return None
    best = min(runs[k][1] for k in cand)
    tied = [k for k in cand if runs[k][1] == best]
    return runs[tied[len(tied) // 2]][0]


def best_lambda(d: PR.SdpData, grid=LAMBDA_GRID):
    """Solve d once per lambda* of the grid (the solver's own GPU solves).
    Returns (label or None, [(lambda, iterations, phase)])."""
    s = Solver(0).load(d).prepare()
    runs = []
    for lam in grid:
        s.set_param(lambdaStar=float(lam))
        r = s.run()
        runs.append((float(lam), int(r.iterations), r.phase))
    s.close()
    return label_from_runs(runs), runs


def generate(seeds, family=scaled_problem):
    """Features and labels (log10 lambda*) over the given problem seeds."""
    F, t, meta = [], [], []
    for sd in seeds:
        d = family(int(sd))
        lam, runs = best_lambda(d)
        print(f"lambda_mlp: {d.name}: label {lam}", flush=True)
        if lam is None:
            continue
        Fb = block_features(d)
        F.append(Fb)
        t.append(np.full(Fb.shape[0], np.log10(lam)))
        meta.append({"seed": int(sd), "name": d.name, "lambda": lam,
                     "iters": {f"{a:g}": b for a, b, _ in runs}})
    return np.concatenate(F), np.concatenate(t), meta


def solve(d: PR.SdpData, model: LambdaModel | None, solver: Solver | None = None):
    """Solve from the predicted per-block lambda* (clipped to 1e-2 .. 1e6);
    if that start does not end in pdOPT, solve again from the default
    lambda* (the fallback).  Returns (result, lambdas or None if the
    fallback was used)."""
    s = solver or Solver(0).load(d).prepare()
    s.set_param(lambdaStar=DEFAULT_LAMBDA)
    if model is None:
        s.set_init(None, None, None)
        return s.run(), None
    lam = np.clip(10.0 ** model.predict_log_lambda(block_features(d)), 1e-2, 1e6)
    X0, Z0, y0 = init_point(d, lam)
    s.set_init(X0, Z0, y0)
    r = s.run()
    if r.phase == "pdOPT":
        return r, lam
    s.set_init(None, None, None)
    return s.run(), None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--train", type=int, default=105,
                    help="training problems of the small family (19 solves each)")
    ap.add_argument("--eval", type=int, default=40, help="held-out problems")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--report", default="")
    ap.add_argument("--large", type=int, default=30,
                    help="training problems of the large family (seeds 2000..)")
    ap.add_argument("--eval-large", type=int, default=8, help="held-out large problems")
    a = ap.parse_args()
    t0 = time.time()
    F, t, meta = generate(range(1000, 1000 + a.train))
    if a.large:
        F2, t2, meta2 = generate(range(2000, 2000 + a.large), large_problem)
        F, t, meta = np.concatenate([F, F2]), np.concatenate([t, t2]), meta + meta2
    print(f"lambda_mlp: {len(meta)} training problems, {len(t)} blocks, "
          f"{(a.train + a.large) * len(LAMBDA_GRID)} solves, {time.time() - t0:.1f}s", flush=True)
    model = LambdaModel(0, seed=7)
    losses = model.fit(F, t, steps=a.steps)
    print(f"lambda_mlp: training loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    model.save(a.out)
    ev = []
    evs = [(sd, scaled_problem) for sd in range(50000, 50000 + a.eval)]
    evs += [(sd, large_problem) for sd in range(60000, 60000 + a.eval_large)]
    for sd, fam in evs:
        d = fam(sd)
        s = Solver(0).load(d).prepare()
        r0, _ = solve(d, None, s)
        r1, lam = solve(d, model, s)
        s.close()
        print(f"lambda_mlp: eval {d.name}: default {r0.phase} {r0.iterations}, model {r1.phase} "
              f"{r1.iterations}{' (fallback)' if lam is None else ''}", flush=True)
        ev.append({"seed": sd, "family": fam.__name__, "default_iters": r0.iterations,
                   "default_phase": r0.phase,
                   "model_iters": r1.iterations, "model_phase": r1.phase,
                   "fallback": lam is None,
                   "obj_rel_diff": abs(r1.primal_obj - r0.primal_obj) / max(1e-12, abs(r0.primal_obj))})
    both = [e for e in ev if e["default_phase"] == e["model_phase"] == "pdOPT"]
    di = float(np.mean([e["default_iters"] for e in both])) if both else None
    mi = float(np.mean([e["model_iters"] for e in both])) if both else None
    rep = {"train_problems": len(meta), "train_blocks": int(len(t)),
           "train_solves": (a.train + a.large) * len(LAMBDA_GRID), "loss_first": losses[0],
           "loss_last": losses[-1], "eval_problems": len(ev),
           "pdOPT_default": sum(e["default_phase"] == "pdOPT" for e in ev),
           "pdOPT_model": sum(e["model_phase"] == "pdOPT" for e in ev),
           "mean_iters_default_where_both_pdOPT": di, "mean_iters_model_where_both_pdOPT": mi,
           "fallbacks": int(sum(e["fallback"] for e in ev)),
           "seconds": time.time() - t0, "eval": ev, "train": meta}
    print(f"lambda_mlp: held-out pdOPT default {rep['pdOPT_default']}/{len(ev)} model "
          f"{rep['pdOPT_model']}/{len(ev)}; mean iterations (both pdOPT) default {di} model {mi} "
          f"(fallbacks {rep['fallbacks']})", flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
