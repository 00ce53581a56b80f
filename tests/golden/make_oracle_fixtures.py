#!/usr/bin/env python3
"""Write the oracle's results on the benchmark workloads as golden fixtures.

The FP64 CPU oracle (oracle/sdpa_oracle.c, a restatement of SDPA 7.4.4's
dense-Schur interior-point loop, sdpa_solve.cpp:10-233) is run here, on the
build machine, on the full-size seeded workloads, and its outcome is kept as
small JSON files so that the GPU tests can compare against it without
re-running minutes of CPU work on the GPU host:

* oracle_solves.json -- for each of
    maxcut_G100_p0.1_seed1          (MaxCut, n = 100, m = 100)
    lovasz_theta_G200_p0.3_seed2    (Lovasz theta, n = 200, m = 6011)
    multiblock_200x[5,50]_lp1000_m5000_seed5
    dense_random_n500_m2000_seed3   (only with --dense: about an hour of CPU)
  the phase, iteration count, objectives, relative gap, feasibility errors
  and the per-iteration trace (mu, objP, objD, alphaP, alphaD, beta, and the
  trace of the Schur matrix at the predictor).
* oracle_batch4096.json -- phase, iterations and objectives of each of the
  4096 batched MaxCut instances (problems.maxcut_batch(4096)), one process
  per instance slice.

Usage: python tests/golden/make_oracle_fixtures.py [--dense] [--only NAME ...]
"""
import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "maxcut_G100_p0.1_seed1": lambda P: P.maxcut(100, 0.1, 1),
    "lovasz_theta_G200_p0.3_seed2": lambda P: P.lovasz(200, 0.3, 2),
    "multiblock_200x[5,50]_lp1000_m5000_seed5": lambda P: P.multi_block(),
    "dense_random_n500_m2000_seed3": lambda P: P.dense_random(500, 2000, 3),
}


def summary(o, seconds):
    return {
        "phase": o.phase, "iterations": int(o.iterations),
        "primal_obj": o.primal_obj, "dual_obj": o.dual_obj,
        "relative_gap": o.relative_gap,
        "primal_error": o.primal_error, "dual_error": o.dual_error,
        "oracle_seconds": seconds,
        # internal convention (SDPA's reversed primal/dual), one row per iteration
        "trace": [{k: t[k] for k in ("mu", "objP", "objD", "alphaP", "alphaD", "beta",
                                     "bmat_trace")} for t in o.trace],
    }


def solve_one(name):
    from oracle import oracle as O
    from sdpa_amd import problems as PR
    d = WORKLOADS[name](PR)
    t0 = time.perf_counter()
    o = O.Problem(d).solve()
    t = time.perf_counter() - t0
    print(f"{name}: {o.phase} in {o.iterations} iterations, {t:.1f} s", flush=True)
    return summary(o, t)


def _batch_slice(args):
    lo, hi = args
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import oracle as O
    from sdpa_amd import problems as PR
    insts = PR.maxcut_batch(4096)
    out = []
    for t in range(lo, hi):
        o = O.Problem(insts[t]).solve()
        out.append([o.phase, int(o.iterations), float("%.13g" % o.primal_obj),
                    float("%.13g" % o.dual_obj)])
    return out


def batch4096(procs):
    n = 4096
    step = (n + 4 * procs - 1) // (4 * procs)
    slices = [(lo, min(n, lo + step)) for lo in range(0, n, step)]
    t0 = time.perf_counter()
    with Pool(procs) as pool:
        parts = pool.map(_batch_slice, slices)
    rows = [r for p in parts for r in p]
    print(f"batch4096: {len(rows)} instances, {time.perf_counter() - t0:.1f} s", flush=True)
    return {"workload": "maxcut_batch(4096): n ~ U{8..64}, p = 0.2, seed 4, trace row on every "
                        "fourth",
            "phase": [r[0] for r in rows], "iterations": [r[1] for r in rows],
            "primal_obj": [r[2] for r in rows], "dual_obj": [r[3] for r in rows]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dense", action="store_true", help="also the dense n=500, m=2000 workload")
    ap.add_argument("--only", nargs="*", default=None)
    ap.add_argument("--procs", type=int, default=8)
    args = ap.parse_args()
    path = os.path.join(HERE, "oracle_solves.json")
    have = json.load(open(path)) if os.path.exists(path) else {}
    names = args.only if args.only else [n for n in WORKLOADS
                                          if args.dense or not n.startswith("dense")]
    for name in names:
        if name == "batch4096":
            res = batch4096(args.procs)
            with open(os.path.join(HERE, "oracle_batch4096.json"), "w") as f:
                json.dump(res, f, separators=(",", ":"))
            continue
        have[name] = solve_one(name)
        with open(path, "w") as f:
            json.dump(have, f, indent=1)


if __name__ == "__main__":
    main()
