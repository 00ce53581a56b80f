#!/usr/bin/env python3
"""Benchmark: time to a 1e-7 gap on the Lovasz-theta SDP of a seeded
G(200, 0.3) graph (seed 2: m = 6011 constraints, one 200x200 block, dense
6011 x 6011 Schur matrix), FP64 GPU path, one MI355X.

A "step" is one complete run of SDPA::initializeSolve + SDPA::solve with the
default parameters (epsilonStar = epsilonDash = 1e-7) on the problem held in
host memory by the solver (as after SDPA::inputElement; file reading is not
timed): the upload to HBM, the structural analysis (changeToDense, formula
choice, Schur decision) and the interior-point loop to pdOPT.  `value` is the
median over the timed steps.  The JSON line also carries:
  - "solve_only": the same loop on the already-uploaded problem;
  - per-phase device times of one extra solve run with HIP-event timing;
  - the F3 bMat assembly bandwidth and the Schur Cholesky's rate;
  - "roofline": the dominant kernel of the Schur factorization (the one of
    its two main kernels with the larger device time: the single-launch panel
    factorization or the wide trailing update; the other in
    "roofline_other"): algorithmic flops per launch over its average launch
    duration, both from HIP events recorded on the stream it runs on;
    "traffic" = HBM bytes per launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a device-copy calibration), or null
    when that summary is absent;
  - "cpu_baseline": one complete solve of the same instance by the FP64 CPU
    oracle (OpenMP over the host threads), same stopping rule;
  - "mixed_vs_fp64": the same workload with the mixed-precision Schur path;
  - "configs": time-to-1e-7 of the other single-GPU workloads (MaxCut
    n = 100, dense-data n = 500 / m = 2000, the 200-block problem, and a
    problem whose Schur matrix takes the sparse branch: 100 blocks of 64 local
    constraints plus 8 coupling ones, m = 6408);
  - "batched": 4096 small MaxCut SDPs in one persistent launch (solves/s,
    phase counts, gap histogram, its own CPU-oracle sample);
  - "lambda_star": the 200-block multi-block SDP (m = 5000) solved from the
    default lambda* = 100 and from the learned per-block lambda* (iterations,
    seconds per solve including feature extraction and prediction).

Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config lovasz|maxcut|dense|multiblock]
  python bench.py --no-lambda --no-configs  # the headline only (kernel profiles)
  python bench.py --copy-calibration        # one 256 MiB device copy (for PMC scaling)
  python bench.py --pmc-summary FETCH_DIR WRITE_DIR panel|trailing   # traffic summary
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
METRIC = "time-to-1e-7 gap; bMat assembly GB/s & MFMA util; mixed vs FP64 GPU, m=6000"
# committed PMC summaries (FETCH_SIZE + WRITE_SIZE per launch) of the Schur
# Cholesky's two main kernels on the headline workload
TRAFFIC_SUMMARY = {
    "panel": os.path.join(ROOT, "profiles", "k_panel_rows_lovasz_m6011_traffic.txt"),
    "trailing": os.path.join(ROOT, "profiles", "k_syrk_update_trailing_lovasz_m6011_traffic.txt"),
}
KERNEL_NAME = {
    "panel": "k_panel_rows<1> (Schur Cholesky: one 256-column panel per launch, FP64 MFMA)",
    "trailing": "k_syrk_update<2,4,3> (Schur Cholesky wide trailing update, FP64 MFMA)",
}
PMC_MATCH = {"panel": "k_panel_rows<1>", "trailing": "k_syrk_update<2, 4, 3>",
             "f3cols": "k_f3_cols", "f3multi": "k_f3_multi"}
COPY_BYTES = 256 << 20

CONFIGS = {
    "lovasz": ("lovasz_theta_G200_p0.3_seed2", lambda P: P.lovasz(200, 0.3, 2)),
    "maxcut": ("maxcut_G100_p0.1_seed1", lambda P: P.maxcut(100, 0.1, 1)),
    "dense": ("dense_random_n500_m2000_seed3", lambda P: P.dense_random(500, 2000, 3)),
    "multiblock": ("multiblock_200x[5,50]_lp1000_m5000_seed5", lambda P: P.multi_block()),
    # the sparse-Schur branch: 100 blocks of 64 local constraints and 8
    # constraints touching every block (an arrow-shaped Schur pattern)
    "sparse": ("block_arrow_100x12_local64_couple8_m6408_seed31",
               lambda P: P.block_separated(n_blocks=100, per_block=64, couple=8, seed=31)),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(d, name):
    """One complete solve of the same instance by the FP64 oracle (OpenMP,
    all host threads it is given), initializeSolve included."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    P = O.Problem(d)
    out = P.solve()
    t = time.perf_counter() - t0
    return {"value": t, "unit": "s", "cores": int(O.lib().orc_num_threads()), "kind": "port",
            "iterations": int(out.iterations), "phase": out.phase,
            "primal_obj": out.primal_obj, "dual_obj": out.dual_obj,
            "sample": f"one complete solve of {name} by the FP64 C oracle with OpenMP "
                      f"({out.iterations} iterations to {out.phase}), problem set-up included"}


def timed_solves(s, steps, warmup, full, label):
    """Median wall time of `steps` solves after `warmup` untimed ones.  full:
    initializeSolve + solve (upload and analysis included), else solve only."""
    if not full:
        s.prepare()
    r = None
    for w in range(warmup):
        r = s.solve() if full else s.run()
        log(f"bench: {label} warmup {w}: phase={r.phase} it={r.iterations}")
    times = []
    for k in range(max(1, steps)):
        t0 = time.perf_counter()
        r = s.solve() if full else s.run()  # returns after the final device->host copy
        times.append(time.perf_counter() - t0)
        log(f"bench: {label} step {k}: phase={r.phase} it={r.iterations} t={times[-1]:.4f}s "
            f"gap={r.relative_gap:.2e}")
    return float(np.median(times)), r


def other_config(key, steps):
    """time-to-1e-7 (initializeSolve + solve, median of `steps`) of another
    single-GPU workload, with its per-phase device times."""
    import sdpa_amd
    from sdpa_amd import problems as PR
    name, gen = CONFIGS[key]
    d = gen(PR)
    s = sdpa_amd.Solver(0).load(d)
    val, r = timed_solves(s, steps, 1, True, name)
    s.enable_timing(True)
    rt = s.run()
    pt = s.phase_times()
    s.enable_timing(False)
    split = None
    if key == "dense":
        # the split-precision path on the same workload: F1/F2 products in
        # FP32 on MFMA, the Schur matrix factored in FP32 with FP64
        # refinement, FP64 from the first refinement stall on
        s.set_param(precision=1)
        st0 = s.mixed_stats()
        mval, rm = timed_solves(s, steps, 1, True, name + " split precision")
        st1 = s.mixed_stats()
        s.enable_timing(True)
        rmt = s.run()
        mpt = s.phase_times()
        s.enable_timing(False)
        s.set_param(precision=0)
        mits = max(1, rmt.iterations)
        fl = d.m * 4.0 * 500 ** 3 + d.m * (d.m + 1.0) * 500 ** 2
        # the assembly runs in FP32 while the factorization is mixed: the
        # iterations before the first stall
        nrun = steps + 1
        mixed_its = (st1["factors"] - st0["factors"]) / nrun
        split = {"time_to_1e7_s": mval, "fp64_over_split": val / mval, "phase": rm.phase,
                 "iterations": rm.iterations, "relative_gap": rm.relative_gap,
                 "primal_error": rm.primal_error, "dual_error": rm.dual_error,
                 "objective_rel_diff_vs_fp64": max(abs(rm.primal_obj - r.primal_obj) / abs(r.primal_obj),
                                                   abs(rm.dual_obj - r.dual_obj) / abs(r.dual_obj)),
                 "fp32_iterations_per_solve": mixed_its,
                 "phase_ms_per_iteration": {k: v / mits for k, v in mpt.items()}}
    s.close()
    its = max(1, rt.iterations)
    sparse = None
    if key == "sparse":
        g = sdpa_amd.GpuContext(0).load(d)
        info = g.info
        g.close()
        sparse = {"schur_dense": bool(info[9]), "tile_factorization": bool(info[12]),
                  "stored_tiles": int(info[13]), "tiles_per_side": int(info[14]),
                  "levels": int(info[15]),
                  "stored_fraction": float(info[13]) / max(1.0, info[14] * (info[14] + 1) / 2.0),
                  "dense_factorization_flops": d.m ** 3 / 3.0}
    out = {"workload": name, "m": d.m, "time_to_1e7_s": val, "phase": r.phase,
           "iterations": r.iterations, "relative_gap": r.relative_gap,
           "primal_error": r.primal_error, "dual_error": r.dual_error,
           "primal_obj": r.primal_obj, "dual_obj": r.dual_obj,
           "phase_ms_per_iteration": {k: v / its for k, v in pt.items()}}
    if key == "dense":
        n, m = 500, d.m
        fl = m * 4.0 * n ** 3 + m * (m + 1.0) * n ** 2  # the reference F1's useful flops
        ms = pt["bmat"] / its
        out["bmat_f1"] = {"flops": fl, "ms": ms, "TFLOPs": fl / (ms * 1e-3) / 1e12,
                          "frac_fp64_peak": fl / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        out["split_precision"] = split
        # the assembly alone in both precisions (HIP-event phase timing, the
        # mean of 3 assemblies at the identity start X = Z = 100 I)
        g = sdpa_amd.GpuContext(0).load(d)
        g.identity_start(100.0)
        asm = {}
        for prec in (0, 1):
            g.timing(1)
            g.make_bmat(prec)  # untimed: the FP32 pool is made by the first FP32 assembly
            g.timing(1)
            for _ in range(3):
                g.make_bmat(prec)
            t = g.timing(0)
            ams = t[0] / max(1.0, t[1])
            peak = FP64_PEAK_TFLOPS if prec == 0 else FP32_PEAK_TFLOPS
            asm["fp64" if prec == 0 else "fp32"] = {
                "ms": ams, "TFLOPs": fl / (ams * 1e-3) / 1e12,
                "frac_peak": fl / (ams * 1e-3) / 1e12 / peak, "peak": peak}
        g.close()
        out["bmat_f1_by_precision"] = asm
    if sparse is not None:
        out["sparse_branch"] = sparse
    del d
    return out


def lambda_star(steps):
    """Default lambda* against the learned per-block lambda* on the 200-block
    multi-block SDP (200 blocks of 5-50, LP 1000, m = 5000, seed 5): best of
    `steps` timed solves each after one untimed, through lambda_mlp.solve (the
    learned start falls back to the default if it does not reach pdOPT)."""
    import sdpa_amd
    from sdpa_amd import lambda_mlp as LM
    from sdpa_amd import problems as PR
    if not os.path.exists(LM.DEFAULT_WEIGHTS):
        return {"value": None, "error": "no trained weights"}
    d = PR.multi_block()
    model = LM.LambdaModel.load()
    s = sdpa_amd.Solver(0).load(d).prepare()
    out = {"workload": "multiblock_200x[5,50]_lp1000_m5000_seed5"}
    for name, mdl in (("default", None), ("learned", model)):
        LM.solve(d, mdl, s)
        best, r, lam = None, None, None
        for _ in range(max(1, steps)):
            t0 = time.perf_counter()
            r, lam = LM.solve(d, mdl, s)
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        out[name] = {"seconds": best, "iterations": int(r.iterations), "phase": r.phase,
                     "relative_gap": r.relative_gap, "primal_obj": r.primal_obj,
                     "fallback": bool(mdl is not None and lam is None)}
        if mdl is not None and lam is not None:
            out[name]["lambda_range"] = [float(lam.min()), float(lam.max())]
    s.close()
    out["speedup"] = out["default"]["seconds"] / out["learned"]["seconds"]
    return out


def batched(steps, warmup, with_cpu):
    """The batched workload: 4096 independent seeded MaxCut SDPs (n ~ U{8..64},
    p = 0.2, a trace row on every fourth), all solved by one persistent launch
    (sdpa_batch_solve).  Reports solves/s over the kernel time (and over the
    whole call, which includes the upload of the packed instances and the
    result copy), the phase counts, the iteration histogram and the gap
    distribution.  The launch configuration is the autotuner's
    (sdpa_batch_autotune: work-group size x heavy-pair threshold x queue
    order, timed on a 512-instance sample); the same batch is also solved with
    the FP32 Schur factorization and FP64 refinement ("mixed")."""
    from sdpa_amd import batch as B
    from sdpa_amd import problems as PR
    insts = PR.maxcut_batch(4096)
    P = B.Packed(insts)
    t0 = time.perf_counter()
    best, table = B.autotune(insts, packed=P, sample=512)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        B.solve_cfg(insts, packed=P, cfg=best)
    kms, wall = [], []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        r = B.solve_cfg(insts, packed=P, cfg=best)
        wall.append(time.perf_counter() - t0)
        kms.append(r.kernel_ms)
    k = float(np.mean(kms)) * 1e-3
    w = float(np.mean(wall))
    phases = {}
    for ph in r.phase:
        phases[ph] = phases.get(ph, 0) + 1
    gaps = r.relative_gap
    mcfg = B.config(best.threads, best.heavy, best.order, precision=1)
    B.solve_cfg(insts, packed=P, cfg=mcfg)
    mk = []
    for _ in range(max(1, steps)):
        rm = B.solve_cfg(insts, packed=P, cfg=mcfg)
        mk.append(rm.kernel_ms)
    mkm = float(np.mean(mk)) * 1e-3
    mph = {}
    for ph in rm.phase:
        mph[ph] = mph.get(ph, 0) + 1
    both = np.array([a == "pdOPT" and b == "pdOPT" for a, b in zip(r.phase, rm.phase)])
    mixed = {"kernel_ms": mkm * 1e3, "value": len(insts) / mkm, "phases": mph,
             "fp64_over_mixed": mkm and k / mkm,
             "max_objective_rel_diff_vs_fp64": float(np.max(
                 np.abs(rm.primal_obj[both] - r.primal_obj[both]) /
                 np.maximum(np.abs(r.primal_obj[both]), 1e-8))) if both.any() else None,
             "schur_factorizations": {"fp32": int(rm.stats[:, 0].sum()),
                                      "fp64": int(rm.stats[:, 1].sum())},
             "refinement_steps_per_fp32_solve": float(rm.stats[:, 2].sum() /
                                                      max(1.0, 2.0 * rm.stats[:, 0].sum())),
             "refinement_stalls": int(rm.stats[:, 3].sum())}
    out = {"workload": "maxcut_batch_4096_n8-64_p0.2_seed4 (trace row on every fourth)",
           "metric": "solves/s", "value": len(insts) / k, "wall_solves_per_s": len(insts) / w,
           "kernel_ms": k * 1e3, "config": best.as_dict(), "autotune_s": tune_s,
           "autotune_table": [{"threads": int(a), "heavy": int(b), "order": int(c), "ms": float(d)}
                              for a, b, c, d in table],
           "mixed": mixed,
           "phases": phases,
           "iterations": {int(i): int(c) for i, c in enumerate(np.bincount(r.iterations)) if c},
           "gap_histogram": {"<=1e-8": int((gaps <= 1e-8).sum()),
                             "(1e-8,1e-7]": int(((gaps > 1e-8) & (gaps <= 1e-7)).sum()),
                             ">1e-7": int((gaps > 1e-7).sum())},
           "max_feasibility_error": float(max(r.primal_error.max(), r.dual_error.max()))}
    if with_cpu:
        # the FP64 oracle on the first 48 instances, one at a time
        from oracle import oracle as O
        sample = insts[:48]
        t0 = time.perf_counter()
        for d in sample:
            O.Problem(d).solve()
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(sample) / t, "unit": "solves/s",
                               "cores": int(O.lib().orc_num_threads()), "kind": "port",
                               "sample": f"the first {len(sample)} of the 4096 instances solved one "
                                         f"after another by the FP64 C oracle ({t:.2f} s)"}
    return out


def read_traffic(which):
    """Traffic line of the committed PMC summary, if present."""
    try:
        with open(TRAFFIC_SUMMARY[which]) as f:
            for line in f:
                mm = re.match(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", line.strip())
                if mm:
                    return float(mm.group(1))
    except OSError:
        pass
    return None


def copy_calibration():
    """One device-to-device copy of COPY_BYTES through the HIP runtime (a
    blit kernel), so that a PMC run can scale FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"bench: copied {COPY_BYTES} bytes device to device")


def _pmc_rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for fn in files:
        with open(fn) as f:
            rows += list(csv.DictReader(f))
    return rows


def _per_dispatch(rows, counter):
    """{dispatch id: (kernel name, value)} for one counter."""
    out = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        did = r.get("Dispatch_Id") or r.get("Correlation_Id")
        name = r.get("Kernel_Name", "")
        v = float(r.get("Counter_Value", 0.0))
        if did in out:
            out[did] = (name, out[did][1] + v)
        else:
            out[did] = (name, v)
    return out


def pmc_summary(fetch_dir, write_dir, which):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of one of the Schur
    Cholesky's main kernels ("panel" or "trailing"), scaled by what the same
    counters read for the calibration copy."""
    lines = []
    res = {}
    name = PMC_MATCH[which]
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        disp = _per_dispatch(_pmc_rows(d), counter)
        dom = [v for (n, v) in disp.values() if name in n]
        cpy = [v for (n, v) in disp.values() if "copyBuffer" in n]
        if not dom or not cpy:
            raise SystemExit(f"{counter}: dominant dispatches {len(dom)}, copy dispatches {len(cpy)}")
        copy_kib = max(cpy)  # the 256 MiB calibration copy is the largest blit
        expect_kib = COPY_BYTES / 1024.0
        factor = expect_kib / copy_kib
        mean_kib = float(np.mean(dom))
        res[counter] = mean_kib * 1024.0 * factor
        lines.append(f"{counter}: {len(dom)} dispatches of {name}, mean {mean_kib:.1f} KiB/dispatch; "
                     f"calibration copy of {COPY_BYTES} B reads {copy_kib:.1f} KiB "
                     f"(expected {expect_kib:.1f}) -> scale {factor:.4f}; "
                     f"scaled {res[counter] / 1e6:.2f} MB/dispatch")
    lines.append(f"traffic_bytes_per_launch: {res['FETCH_SIZE'] + res['WRITE_SIZE']:.6e}")
    print("\n".join(lines))


def pmc_counters(d, which):
    """Mean per dispatch of every counter in a PMC run, over the dispatches of
    one kernel (a PMC_MATCH key)."""
    rows = _pmc_rows(d)
    name = PMC_MATCH[which]
    counters = sorted({r.get("Counter_Name") for r in rows if name in r.get("Kernel_Name", "")})
    for counter in counters:
        vals = [v for (n, v) in _per_dispatch(rows, counter).values() if name in n]
        print(f"{counter}: {len(vals)} dispatches of {name}, mean {float(np.mean(vals)):.6e}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="lovasz", choices=list(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lambda", action="store_true",
                    help="skip the learned-lambda* comparison (kernel profiles of the headline)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other single-GPU workloads")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, one 256 MiB device copy (PMC scaling)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "KERNEL"),
                    help="KERNEL: " + ", ".join(PMC_MATCH))
    ap.add_argument("--pmc-counters", nargs=2, metavar=("DIR", "KERNEL"),
                    help="mean per dispatch of each counter of a PMC run, for KERNEL")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.pmc_counters:
        pmc_counters(*args.pmc_counters)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the design notes describe how the Schur factorization would shard).",
              file=sys.stderr)
        sys.exit(2)

    import sdpa_amd
    from sdpa_amd import problems as PR

    name, gen = CONFIGS[args.config]
    d = gen(PR)
    n_sdp = int(d.block_struct[0])
    log(f"bench: config {name}: m={d.m} blocks={list(d.block_struct)[:4]} nnz={d.nnz}")
    s = sdpa_amd.Solver(0).load(d)
    value, r = timed_solves(s, args.steps, args.warmup, True, "fp64")
    solve_only, _ = timed_solves(s, args.steps, 1, False, "fp64 solve-only")
    # one extra solve with per-phase HIP-event timing (synchronizing, not in `value`)
    s.enable_timing(True)
    rt = s.run()
    pt = s.phase_times()
    ktime = s.kernel_timing()
    s.enable_timing(False)
    log(f"bench: timed solve: phase={rt.phase} it={rt.iterations}; kernel timing {ktime}")
    # the same workload with the mixed-precision Schur factorization (FP32
    # factor + FP64 refinement, FP64 after the first stall), same step count
    s.set_param(precision=1)
    st0 = s.mixed_stats()
    mval, rm = timed_solves(s, args.steps, args.warmup, True, "mixed")
    st1 = s.mixed_stats()
    s.enable_timing(True)
    s.run()
    mpt = s.phase_times()
    s.enable_timing(False)
    s.set_param(precision=0)
    s.close()
    nrun = max(1, args.steps) + args.warmup
    mixed = {
        "value": mval, "unit": "s",
        "fp64_over_mixed": value / mval,
        "phase": rm.phase, "iterations": rm.iterations, "relative_gap": rm.relative_gap,
        "primal_error": rm.primal_error, "dual_error": rm.dual_error,
        "objective_rel_diff_vs_fp64": max(abs(rm.primal_obj - rt.primal_obj) / abs(rt.primal_obj),
                                          abs(rm.dual_obj - rt.dual_obj) / abs(rt.dual_obj)),
        "per_solve": {k: (st1[k] - st0[k]) / nrun for k in
                      ("factors", "fallbacks", "solves", "refine_steps", "stalls")},
        "phase_ms_per_iteration": {k: v / max(1, rm.iterations) for k, v in mpt.items()},
    }
    configs = None
    if not args.no_configs:
        configs = {}
        for key in ("maxcut", "dense", "multiblock", "sparse"):
            if key == args.config:
                continue
            log(f"bench: config {key}")
            try:
                configs[key] = other_config(key, min(3, max(1, args.steps)))
            except Exception as e:  # reported beside the headline, never part of it
                configs[key] = {"error": str(e)}
    log("bench: batched workload")
    batch = batched(min(args.steps, 5), min(args.warmup, 2), not args.no_cpu_baseline)
    log(f"bench: batched {batch['value']:.0f} solves/s ({batch['kernel_ms']:.1f} ms kernel)")
    lam = None
    if not args.no_lambda:
        log("bench: learned lambda* on the 200-block configuration")
        try:
            lam = lambda_star(min(args.steps, 3))
        except Exception as e:  # reported beside the headline, never part of it
            lam = {"value": None, "error": str(e)}
        log(f"bench: lambda* {json.dumps(lam)[:200]}")
    its = rt.iterations
    m = d.m
    # F3 bytes per assembly (lower triangle written): 8*T + 16*sum(nnz) + 16*n^2
    bmat_bytes = 8 * m * (m + 1) / 2 + 16 * d.nnz + 16 * n_sdp * n_sdp
    bmat_ms = pt["bmat"] / max(1, its)
    chol_ms = pt["chol"] / max(1, its)
    # roofline of the factorization's dominant kernel (the larger total device
    # time of its two main kernels), the other beside it
    roofs = {}
    for key, (k_ms, k_n, k_flops) in ktime.items():
        if k_n > 0 and k_ms > 0:
            avg_ms = k_ms / k_n
            fl = k_flops / k_n
            ach = fl / (avg_ms * 1e-3) / 1e12
            roofs[key] = {"kernel": KERNEL_NAME[key], "bound": "mfma", "achieved": ach,
                          "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": ach / FP64_PEAK_TFLOPS, "traffic": read_traffic(key),
                          "flops_per_launch": fl, "avg_launch_us": avg_ms * 1e3, "launches": k_n,
                          "total_ms_per_solve": k_ms,
                          "timing": "HIP events on the kernel's stream around each launch, "
                                    "one timed solve"}
    dom = max(roofs, key=lambda k: roofs[k]["total_ms_per_solve"]) if roofs else None
    roof = roofs.get(dom)
    other_roof = {k: v for k, v in roofs.items() if k != dom}
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, sdpa_amd/problems.py)",
        "config": {"workload": name, "m": m, "n": n_sdp, "iterations": its,
                   "phase": rt.phase, "relative_gap": rt.relative_gap,
                   "primal_error": rt.primal_error, "dual_error": rt.dual_error,
                   "primal_obj": rt.primal_obj, "dual_obj": rt.dual_obj},
        "phase_ms_per_iteration": {k: v / max(1, its) for k, v in pt.items()},
        "bmat_f3": {"bytes": bmat_bytes, "ms": bmat_ms, "GBps": bmat_gbs(bmat_bytes, bmat_ms),
                    "frac_hbm": bmat_gbs(bmat_bytes, bmat_ms) / HBM_PEAK_GBS},
        "cholesky": {"flops": m ** 3 / 3.0, "ms": chol_ms,
                     "TFLOPs": m ** 3 / 3.0 / (chol_ms * 1e-3) / 1e12 if chol_ms > 0 else None},
        "solve_only": {"value": solve_only, "unit": "s",
                       "note": "SDPA::solve on the uploaded problem (initializeSolve not timed)"},
        "mixed_vs_fp64": mixed,
        "configs": configs,
        "batched": batch,
        "lambda_star": lam,
        "roofline": roof,
        "roofline_other": other_roof,
    }
    if not args.no_cpu_baseline:
        try:
            cb = cpu_baseline(d, name)
            line["cpu_baseline"] = cb
            # the GPU results' objective errors against the FP64 oracle's
            for key, rr in (("fp64", rt), ("mixed", rm)):
                line.setdefault("error_vs_oracle", {})[key] = {
                    "primal_obj_rel": abs(rr.primal_obj - cb["primal_obj"]) / abs(cb["primal_obj"]),
                    "dual_obj_rel": abs(rr.dual_obj - cb["dual_obj"]) / abs(cb["dual_obj"]),
                    "iterations": [int(rr.iterations), cb["iterations"]]}
        except Exception as e:  # the baseline is reported, never the target
            line["cpu_baseline"] = {"value": None, "error": str(e)}
    if args.copy_calibration:
        copy_calibration()
    print(json.dumps(line), flush=True)


def bmat_gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9 if ms > 0 else None


if __name__ == "__main__":
    main()
