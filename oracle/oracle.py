"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
routines compiled under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsdparef.so")

PHASES = ["noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF", "pFEAS_dINF", "pINF_dFEAS",
          "pdOPT", "pUNBD", "dUNBD"]
# Phase::reverse, sdpa_parts.cpp:1244-1264
REVERSE = {"pFEAS": "dFEAS", "dFEAS": "pFEAS", "pFEAS_dINF": "pINF_dFEAS",
           "pINF_dFEAS": "pFEAS_dINF", "pUNBD": "dUNBD", "dUNBD": "pUNBD"}


class Param(C.Structure):
    _fields_ = [("maxIteration", C.c_int), ("epsilonStar", C.c_double), ("lambdaStar", C.c_double),
                ("omegaStar", C.c_double), ("lowerBound", C.c_double), ("upperBound", C.c_double),
                ("betaStar", C.c_double), ("betaBar", C.c_double), ("gammaStar", C.c_double),
                ("epsilonDash", C.c_double), ("kappa", C.c_double)]


class Trace(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta", "bmat_trace")]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("phase", C.c_int), ("cholFail", C.c_int),
                ("stepFail", C.c_int), ("invFail", C.c_int)] + \
               [(n, C.c_double) for n in ("mu", "objValPrimal", "objValDual", "normPrimalVec",
                                          "normDualMat", "time_total", "time_bmat", "time_chol",
                                          "time_solve", "time_eig", "time_inv")]


_lib = None
_ref = None
_P = C.c_void_p
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_problem_create.restype = _P
        L.orc_problem_create.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_long, _ip, _ip, _ip, _ip, _dp]
        L.orc_problem_prepare.argtypes = [_P, C.c_double]
        L.orc_problem_free.argtypes = [_P]
        L.orc_packed_size.restype = C.c_long
        L.orc_packed_size.argtypes = [_P]
        L.orc_n_dim.argtypes = [_P]
        L.orc_schur_dense.argtypes = [_P]
        L.orc_formula.argtypes = [_P, C.c_int, C.c_int]
        L.orc_formula_counts.argtypes = [_P, _ip]
        L.orc_chol_adjust.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_chol_solve.argtypes = [C.c_int, _dp, C.c_int, _dp]
        L.orc_inv_chol.argtypes = [C.c_int, _dp, _dp, C.c_void_p, C.c_void_p]
        L.orc_tridiag_ql.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int]
        L.orc_sym_min_eig.restype = C.c_double
        L.orc_sym_min_eig.argtypes = [C.c_int, _dp]
        L.orc_lanczos_min.restype = C.c_double
        L.orc_lanczos_min.argtypes = [C.c_int, _dp, _dp]
        L.orc_min_eig_packed.restype = C.c_double
        L.orc_min_eig_packed.argtypes = [_P, _dp, _dp]
        L.orc_make_bmat.argtypes = [_P, _dp, _dp, _dp]
        L.orc_direction.argtypes = [_P, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                    C.c_double, _dp]
        L.orc_default_param.argtypes = [C.POINTER(Param)]
        L.orc_solve.argtypes = [_P, C.POINTER(Param), C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(Trace), C.POINTER(Result), C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_int]
        assert L.orc_sizeof_param() == C.sizeof(Param)
        assert L.orc_sizeof_trace() == C.sizeof(Trace)
        assert L.orc_sizeof_result() == C.sizeof(Result)
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """Reference routines compiled from the SDPA sources (build container only)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_chol_adjust.argtypes = [C.c_int, _dp, C.c_int]
        R.ref_lanczos_min.restype = C.c_double
        R.ref_lanczos_min.argtypes = [C.c_int, _dp, _dp]
        R.ref_small_min_eig.restype = C.c_double
        R.ref_small_min_eig.argtypes = [C.c_int, _dp, _dp]
        R.ref_inv_chol.argtypes = [C.c_int, _dp, _dp, _dp]
        _ref = R
    return _ref


def default_param(**kw) -> Param:
    p = Param()
    lib().orc_default_param(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Problem:
    """An SDP in the oracle, built from an SdpData (sdpa_amd.problems)."""

    def __init__(self, d, kappa: float = 12.0):
        L = lib()
        self.d = d
        self._arrs = [np.ascontiguousarray(a) for a in (d.block_struct.astype(np.int32), d.c,
                                                          d.k, d.l, d.i, d.j, d.v)]
        bs, c, k, l, i, j, v = self._arrs
        self.h = L.orc_problem_create(d.m, len(bs), bs, c, len(k), k, l, i, j, v)
        L.orc_problem_prepare(self.h, kappa)
        self.m = d.m
        self.total = L.orc_packed_size(self.h)
        self.n_dim = L.orc_n_dim(self.h)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_problem_free(self.h)
        except Exception:
            pass

    @property
    def schur_dense(self) -> bool:
        return bool(lib().orc_schur_dense(self.h))

    def formula_counts(self):
        out = np.zeros(3, np.int32)
        lib().orc_formula_counts(self.h, out)
        return out

    def make_bmat(self, X, Zinv):
        B = np.zeros(self.m * self.m)
        lib().orc_make_bmat(self.h, np.ascontiguousarray(X, np.float64),
                            np.ascontiguousarray(Zinv, np.float64), B)
        return B.reshape(self.m, self.m, order="F")

    def min_eig(self, Lp, Dp) -> float:
        return lib().orc_min_eig_packed(self.h, np.ascontiguousarray(Lp, np.float64),
                                        np.ascontiguousarray(Dp, np.float64))

    def direction(self, X, y, Z, infeasible: bool, mu: float, beta_pred: float,
                  beta_corr: float):
        """Predictor and corrector Newton directions from (X, y, Z) (internal
        convention, packed): two dicts of g, dy, Dx, Dz."""
        m, T = self.m, self.total
        out = np.zeros(2 * (2 * m + 2 * T))
        rc = lib().orc_direction(self.h, np.ascontiguousarray(X, np.float64),
                                 np.ascontiguousarray(y, np.float64),
                                 np.ascontiguousarray(Z, np.float64), int(bool(infeasible)),
                                 float(mu), float(beta_pred), float(beta_corr), out)
        if rc != 0:
            raise RuntimeError(f"orc_direction failed ({rc})")
        res = []
        for p in range(2):
            o = out[p * (2 * m + 2 * T):(p + 1) * (2 * m + 2 * T)]
            res.append({"g": o[:m].copy(), "dy": o[m:2 * m].copy(),
                        "Dx": o[2 * m:2 * m + T].copy(), "Dz": o[2 * m + T:].copy()})
        return res[0], res[1]

    def solve(self, param: Param | None = None, X0=None, Z0=None, y0=None, max_iter: int = 0):
        L = lib()
        p = param or default_param()
        nmax = (max_iter or p.maxIteration) + 2
        tr = (Trace * nmax)()
        res = Result()
        X = np.zeros(self.total); Z = np.zeros(self.total); y = np.zeros(self.m)
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        L.orc_solve(self.h, C.byref(p), ptr(X0), ptr(Z0), ptr(y0), tr, C.byref(res),
                    X.ctypes.data_as(C.c_void_p), Z.ctypes.data_as(C.c_void_p),
                    y.ctypes.data_as(C.c_void_p), max_iter)
        its = res.iterations
        trace = [{f: getattr(tr[t], f) for f, _ in Trace._fields_} for t in range(its + 1)]
        return SolveOutput(res, trace, X, Z, y)


class SolveOutput:
    """Results in both conventions.  Internal: SDPA's reversed primal/dual;
    user-facing getters follow sdpa_call.cpp:668-747."""

    def __init__(self, res: Result, trace, X, Z, y):
        self.res = res
        self.trace = trace
        self.X_int, self.Z_int, self.y_int = X, Z, y
        self.iterations = res.iterations
        ph = PHASES[res.phase]
        self.phase_internal = ph
        self.phase = REVERSE.get(ph, ph)
        # user convention: primal objective = -objValDual(internal), etc.
        self.primal_obj = -res.objValDual
        self.dual_obj = -res.objValPrimal
        self.primal_error = res.normDualMat
        self.dual_error = res.normPrimalVec
        self.x = -y  # getResultXVec after yVec *= -1 (sdpa_solve.cpp:200-203)
        self.mu = res.mu

    def user_trace(self):
        """printOneIteration columns with REVERSE_PRIMAL_DUAL (sdpa_io.cpp:638-644)."""
        return [(t["mu"], t["thetaD"], t["thetaP"], -t["objD"], -t["objP"], t["alphaD"],
                 t["alphaP"], t["beta"]) for t in self.trace]

    @property
    def relative_gap(self):
        a, b = self.res.objValPrimal, self.res.objValDual
        mean = (abs(a) + abs(b)) / 2.0
        gap = abs(a - b)
        return gap if mean < 1.0 else gap / mean


def chol_adjust(A: np.ndarray):
    A = np.asfortranarray(A, np.float64).copy(order="F")
    flat = A.reshape(-1, order="F")
    flat = np.ascontiguousarray(flat)
    info = lib().orc_chol_adjust(A.shape[0], flat, A.shape[0])
    return info, flat.reshape(A.shape, order="F")


def chol_solve(L: np.ndarray, g: np.ndarray):
    x = np.ascontiguousarray(g, np.float64).copy()
    flat = np.ascontiguousarray(np.asarray(L, np.float64).reshape(-1, order="F"))
    lib().orc_chol_solve(L.shape[0], flat, L.shape[0], x)
    return x


def inv_chol(A: np.ndarray):
    n = A.shape[0]
    a = np.ascontiguousarray(np.asarray(A, np.float64).reshape(-1, order="F"))
    Linv = np.zeros(n * n); Ainv = np.zeros(n * n)
    info = lib().orc_inv_chol(n, a, Linv, Ainv.ctypes.data_as(C.c_void_p), None)
    return info, Linv.reshape(n, n, order="F"), Ainv.reshape(n, n, order="F")


def lanczos_min(L: np.ndarray, D: np.ndarray) -> float:
    n = L.shape[0]
    return lib().orc_lanczos_min(n, np.ascontiguousarray(L.reshape(-1, order="F")),
                                 np.ascontiguousarray(D.reshape(-1, order="F")))


def sym_min_eig(A: np.ndarray) -> float:
    return lib().orc_sym_min_eig(A.shape[0], np.ascontiguousarray(A.reshape(-1, order="F")))


def tridiag_ql(d: np.ndarray, e: np.ndarray):
    n = d.shape[0]
    dd = np.ascontiguousarray(d, np.float64).copy()
    ee = np.zeros(max(n, 1)); ee[: n - 1] = e
    Z = np.zeros(n * n)
    info = lib().orc_tridiag_ql(n, dd, ee, Z, n)
    return info, dd, Z.reshape(n, n, order="F")
