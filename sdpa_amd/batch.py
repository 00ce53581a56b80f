"""Batched small-SDP solver: many independent single-block instances solved by
one persistent launch (``sdpa_batch_solve``, include/sdpa_batch.h).

Each instance is an :class:`~sdpa_amd.problems.SdpData` with one SDP block of
size n <= 64 and m <= 72 constraints.  The device runs SDPA's whole loop per
instance (SDPA::solve, sdpa_solve.cpp:59-233); the results come back in the
user convention of sdpa_call.cpp:668-747 (primal objective = -b.y of the
internal problem, and so on), as :class:`BatchResult` arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import PHASES, GpuError, lib

NMAX, MMAX = 64, 72
# Parameter::setDefaultParameter (sdpa_parts.cpp:1273-1290): maxIteration,
# epsilonStar, epsilonDash, betaStar, betaBar, gammaStar, omegaStar,
# lowerBound, upperBound -- the order of the C ABI's param array
DEFAULT = dict(maxIteration=100, epsilonStar=1.0e-7, epsilonDash=1.0e-7, betaStar=0.1,
               betaBar=0.2, gammaStar=0.9, omegaStar=2.0, lowerBound=-1.0e5, upperBound=1.0e5)
_ORDER = ("maxIteration", "epsilonStar", "epsilonDash", "betaStar", "betaBar", "gammaStar",
          "omegaStar", "lowerBound", "upperBound")
# Phase::reverse (sdpa_parts.cpp:1244-1264) on the phase codes
_REV = {1: 2, 2: 1, 5: 6, 6: 5, 8: 9, 9: 8}

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def _bind():
    L = lib()
    f = L.sdpa_batch_solve
    if f.argtypes is None:
        f.argtypes = [C.c_int, C.c_int, _ip, _ip, _lp, _dp, C.c_long, _lp, _dp, C.c_long, _lp, _ip,
                      C.c_long, _ip, _ip, _dp, C.c_long, _dp, _dp, C.c_int, _dp,
                      C.POINTER(C.c_double)]
    return f


class Config(C.Structure):
    """sdpa_batch_config (include/sdpa_batch.h): launch shape and precision."""
    _fields_ = [("threads", C.c_int), ("heavy", C.c_int), ("order", C.c_int),
                ("precision", C.c_int), ("mu_switch", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def config(threads=256, heavy=64, order=0, precision=0, mu_switch=0.0) -> Config:
    return Config(int(threads), int(heavy), int(order), int(precision), float(mu_switch))


def _bind_cfg():
    L = lib()
    f = L.sdpa_batch_solve_cfg
    if f.argtypes is None:
        f.argtypes = [C.c_int, C.c_int, _ip, _ip, _lp, _dp, C.c_long, _lp, _dp, C.c_long, _lp, _ip,
                      C.c_long, _ip, _ip, _dp, C.c_long, _dp, _dp, C.POINTER(Config), _dp,
                      C.c_void_p, C.POINTER(C.c_double)]
        g = L.sdpa_batch_autotune
        g.argtypes = [C.c_int, C.c_int, _ip, _ip, _lp, _dp, C.c_long, _lp, _dp, C.c_long, _lp, _ip,
                      C.c_long, _ip, _ip, _dp, C.c_long, _dp, _dp, C.c_int, C.c_double, C.c_int,
                      C.POINTER(Config), C.c_void_p, C.c_int]
    return f


class Packed:
    """Instances in the C ABI's concatenated layout (internal convention:
    C = -F0, b = c, A_i = F_i with its upper-triangle entries, 0-based)."""

    def __init__(self, problems):
        cnt = len(problems)
        self.count = cnt
        self.n = np.zeros(cnt, np.int32)
        self.m = np.zeros(cnt, np.int32)
        self.coff = np.zeros(cnt, np.int64)
        self.boff = np.zeros(cnt, np.int64)
        self.pptr = np.zeros(cnt, np.int64)
        Cs, bs, cps, ers, ecs, evs = [], [], [], [], [], []
        co = bo = po = eo = 0
        for t, d in enumerate(problems):
            if d.n_block != 1 or int(d.block_struct[0]) < 1:
                raise ValueError(f"instance {t}: one SDP block required")
            n, m = int(d.block_struct[0]), int(d.m)
            if n > NMAX or m > MMAX:
                raise ValueError(f"instance {t}: n={n} m={m} beyond {NMAX}/{MMAX}")
            r = np.minimum(d.i, d.j).astype(np.int64) - 1
            c = np.maximum(d.i, d.j).astype(np.int64) - 1
            k = d.k.astype(np.int64)
            Cm = np.zeros((n, n))
            f0 = k == 0
            np.add.at(Cm, (r[f0], c[f0]), -d.v[f0])
            off = f0 & (r != c)
            np.add.at(Cm, (c[off], r[off]), -d.v[off])
            # constraint entries, merged per (k, r, c) and ordered by (k, c, r)
            sel = ~f0
            key = (k[sel] * n + c[sel]) * n + r[sel]
            uk, inv = np.unique(key, return_inverse=True)
            vals = np.zeros(len(uk))
            np.add.at(vals, inv, d.v[sel])
            kk = uk // (n * n)
            cc = (uk // n) % n
            rr = uk % n
            ptr = np.searchsorted(kk, np.arange(1, m + 2)) + eo
            self.n[t], self.m[t] = n, m
            self.coff[t], self.boff[t], self.pptr[t] = co, bo, po
            Cs.append(Cm.reshape(-1, order="F"))
            bs.append(np.asarray(d.c, np.float64))
            cps.append(ptr.astype(np.int32))
            ers.append(rr.astype(np.int32))
            ecs.append(cc.astype(np.int32))
            evs.append(vals)
            co += n * n
            bo += m
            po += m + 1
            eo += len(uk)
        self.C = np.ascontiguousarray(np.concatenate(Cs))
        self.b = np.ascontiguousarray(np.concatenate(bs))
        self.cptr = np.ascontiguousarray(np.concatenate(cps))
        self.er = np.ascontiguousarray(np.concatenate(ers)) if ers else np.zeros(0, np.int32)
        self.ec = np.ascontiguousarray(np.concatenate(ecs)) if ecs else np.zeros(0, np.int32)
        self.ev = np.ascontiguousarray(np.concatenate(evs)) if evs else np.zeros(0)


class BatchResult:
    """Per-instance results (user convention), arrays of length count.
    stats (solve_cfg only): per instance the Schur factorizations in FP32 and
    in FP64, the refinement steps and the refinement stalls (count x 4)."""

    def __init__(self, out, kernel_ms, stats=None):
        o = out.reshape(-1, 8)
        ph = o[:, 0].astype(int)
        self.phase_internal = ph
        self.phase_code = np.array([_REV.get(int(p), int(p)) for p in ph])
        self.phase = [PHASES[p] for p in self.phase_code]
        self.iterations = o[:, 1].astype(int)
        self.primal_obj = -o[:, 3]
        self.dual_obj = -o[:, 2]
        self.primal_error = o[:, 5]
        self.dual_error = o[:, 4]
        self.mu = o[:, 6]
        self.failure = o[:, 7].astype(int)
        a, b = o[:, 2], o[:, 3]
        mean = (np.abs(a) + np.abs(b)) / 2.0
        gap = np.abs(a - b)
        self.relative_gap = np.where(mean < 1.0, gap, gap / np.maximum(mean, 1.0))
        self.kernel_ms = kernel_ms
        self.stats = None if stats is None else stats.reshape(-1, 4)


def _bind_noalloc():
    L = lib()
    f = L.sdpa_batch_solve_noalloc
    if f.argtypes is None:
        f.argtypes = [C.c_int, C.c_int, _ip, _ip, _lp, _dp, C.c_long, _lp, _dp, C.c_long, _lp, _ip,
                      C.c_long, _ip, _ip, _dp, C.c_long, _dp, _dp, C.c_int, _dp, C.c_void_p]
        L.sdpa_batch_noalloc_capacity.restype = C.c_long
        L.sdpa_batch_noalloc_capacity.argtypes = []
    return f


def _args(problems, lambdas, packed, params):
    P = packed if packed is not None else Packed(problems)
    lam = np.full(P.count, 100.0) if lambdas is None else np.ascontiguousarray(lambdas, np.float64)
    if lam.shape != (P.count,):
        raise ValueError("one lambda per instance")
    pr = dict(DEFAULT)
    for k, v in params.items():
        if k not in pr:
            raise ValueError(f"unknown parameter {k}")
        pr[k] = v
    return P, lam, np.array([float(pr[k]) for k in _ORDER])


def solve_noalloc(problems, lambdas=None, device: int = 0, threads: int = 256,
                  packed: Packed | None = None, **params) -> BatchResult:
    """The same solve through sdpa_batch_solve_noalloc: no device allocation
    (the library's fixed scratch array holds the launch's buffers; batches
    that do not fit it raise).  kernel_ms of the result is 0."""
    P, lam, par = _args(problems, lambdas, packed, params)
    out = np.zeros(8 * P.count)
    rc = _bind_noalloc()(device, P.count, P.n, P.m, P.coff, P.C, len(P.C), P.boff, P.b, len(P.b),
                         P.pptr, P.cptr, len(P.cptr), P.er, P.ec, P.ev, len(P.ev), lam, par,
                         threads, out, None)
    if rc != 0:
        raise GpuError(f"sdpa_batch_solve_noalloc failed ({rc})")
    return BatchResult(out, 0.0)


def solve(problems, lambdas=None, device: int = 0, threads: int = 256, packed: Packed | None = None,
          **params) -> BatchResult:
    """Solve every instance from X0 = Z0 = lambda_k I, y0 = 0 (lambda_k defaults
    to SDPA's lambdaStar, 100).  Keyword parameters override DEFAULT."""
    P = packed if packed is not None else Packed(problems)
    lam = np.full(P.count, 100.0) if lambdas is None else np.ascontiguousarray(lambdas, np.float64)
    if lam.shape != (P.count,):
        raise ValueError("one lambda per instance")
    pr = dict(DEFAULT)
    for k, v in params.items():
        if k not in pr:
            raise ValueError(f"unknown parameter {k}")
        pr[k] = v
    par = np.array([float(pr[k]) for k in _ORDER])
    out = np.zeros(8 * P.count)
    ms = C.c_double(0.0)
    rc = _bind()(device, P.count, P.n, P.m, P.coff, P.C, len(P.C), P.boff, P.b, len(P.b), P.pptr,
                 P.cptr, len(P.cptr), P.er, P.ec, P.ev, len(P.ev), lam, par, threads, out,
                 C.byref(ms))
    if rc != 0:
        raise GpuError(f"sdpa_batch_solve failed ({rc})")
    return BatchResult(out, ms.value)


def solve_cfg(problems, lambdas=None, device: int = 0, cfg: Config | None = None,
              packed: Packed | None = None, **params) -> BatchResult:
    """solve() under a launch configuration (threads, heavy-pair threshold,
    queue order, Schur precision and mu switch); the result carries the
    per-instance factorization and refinement counts in .stats."""
    P, lam, par = _args(problems, lambdas, packed, params)
    cfg = cfg if cfg is not None else config()
    out = np.zeros(8 * P.count)
    st = np.zeros(4 * P.count)
    ms = C.c_double(0.0)
    rc = _bind_cfg()(device, P.count, P.n, P.m, P.coff, P.C, len(P.C), P.boff, P.b, len(P.b),
                     P.pptr, P.cptr, len(P.cptr), P.er, P.ec, P.ev, len(P.ev), lam, par,
                     C.byref(cfg), out, st.ctypes.data_as(C.c_void_p), C.byref(ms))
    if rc != 0:
        raise GpuError(f"sdpa_batch_solve_cfg failed ({rc})")
    return BatchResult(out, ms.value, st)


def autotune(problems, lambdas=None, device: int = 0, precision: int = 0,
             mu_switch: float = 0.0, sample: int = 512, packed: Packed | None = None,
             **params):
    """The launch configuration search of sdpa_batch_autotune on a sample of
    the batch.  Returns (best Config, table of (threads, heavy, order, ms))."""
    P, lam, par = _args(problems, lambdas, packed, params)
    _bind_cfg()
    best = Config()
    cap = 64
    table = np.zeros(4 * cap)
    rc = lib().sdpa_batch_autotune(device, P.count, P.n, P.m, P.coff, P.C, len(P.C), P.boff, P.b,
                                   len(P.b), P.pptr, P.cptr, len(P.cptr), P.er, P.ec, P.ev,
                                   len(P.ev), lam, par, int(precision), float(mu_switch),
                                   int(sample), C.byref(best), table.ctypes.data_as(C.c_void_p),
                                   cap)
    if rc <= 0:
        raise GpuError(f"sdpa_batch_autotune failed ({rc})")
    return best, table[: 4 * min(rc, cap)].reshape(-1, 4)
