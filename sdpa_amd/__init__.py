"""sdpa_amd -- MI355X (gfx950) Newton-direction hot path of SDPA 7.4.4.

Python access to ``libsdpa_amd.so`` (built in-tree by ``make -C sdpa_amd``):

* :class:`Solver` -- the host interior-point driver (``sdpa_solver_*`` C
  wrapper of class ``sdpa_amd::SDPA``, a mirror of SDPA's ``sdpa_call.h``).
* :class:`GpuContext` -- the kernel-level C ABI of ``include/sdpa_gpu.h``
  (bMat assembly, Schur Cholesky and solve, X/Z inverse Cholesky, step-length
  eigenvalues, direction assembly).

There is no CPU fallback: if the library is missing or cannot be loaded the
import of these classes raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdpa_amd.so")

PHASES = ["noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF", "pFEAS_dINF", "pINF_dFEAS",
          "pdOPT", "pUNBD", "dUNBD"]

# sdpa_gpu_dots codes
DOT = dict(CX=0, BY=1, CDX=2, BDY=3, XDZ=4, DXZ=5, DXDZ=6, XZ=7, TRX=8, TRZ=9, X0Z=10, XZ0=11)
# sdpa_gpu_get/set_buffer ids
BUF = dict(X=0, Z=1, y=2, Zinv=3, invCholX=4, invCholZ=5, Dx=6, Dz=7, dy=8, g=9, B=10,
           pvec=11, dmat=12)

_lib = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class LibraryMissing(RuntimeError):
    pass


def build(force: bool = True) -> str:
    """Compile libsdpa_amd.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    import subprocess
    if force:
        subprocess.check_call(["make", "-s", "-C", HERE, "clean"])
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built; run `make -C sdpa_amd` "
                                 "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        V = C.c_void_p
        L.sdpa_gpu_create.argtypes = [C.c_int, C.POINTER(V)]
        L.sdpa_gpu_destroy.argtypes = [V]
        L.sdpa_gpu_last_error.restype = C.c_char_p
        L.sdpa_gpu_last_error.argtypes = [V]
        L.sdpa_gpu_load_problem.argtypes = [V, C.c_int, C.c_int, _ip, _dp, C.c_long, _ip, _ip, _ip,
                                            _ip, _dp, C.c_double]
        L.sdpa_gpu_problem_info.argtypes = [V, np.ctypeslib.ndpointer(np.int64), C.c_int]
        L.sdpa_gpu_block_size.argtypes = [V, C.c_int]
        for f in ("sdpa_gpu_set_iterate", "sdpa_gpu_get_iterate"):
            getattr(L, f).argtypes = [V, V, V, V]
        L.sdpa_gpu_identity_start.argtypes = [V, C.c_double]
        L.sdpa_gpu_keep_initial_point.argtypes = [V]
        L.sdpa_gpu_inv_chol.argtypes = [V]
        L.sdpa_gpu_make_bmat.argtypes = [V, C.c_int]
        L.sdpa_gpu_factor_bmat.argtypes = [V, C.c_int]
        L.sdpa_gpu_solve_bmat.argtypes = [V, C.c_int]
        L.sdpa_gpu_min_eig.argtypes = [V, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sdpa_gpu_residuals.argtypes = [V, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sdpa_gpu_make_gvec.argtypes = [V, C.c_int, C.c_double, C.c_int]
        L.sdpa_gpu_make_direction.argtypes = [V, C.c_int]
        L.sdpa_gpu_dots.argtypes = [V, C.c_int, _ip, _dp]
        L.sdpa_gpu_update.argtypes = [V, C.c_double, C.c_double]
        L.sdpa_gpu_min_eig_dots.argtypes = [V, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.c_int, _ip, _dp]
        L.sdpa_gpu_update_end.argtypes = [V, C.c_int, _ip, _dp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
        L.sdpa_gpu_get_buffer.argtypes = [V, C.c_int, _dp, C.c_long]
        L.sdpa_gpu_set_buffer.argtypes = [V, C.c_int, _dp, C.c_long]
        L.sdpa_gpu_timing.argtypes = [V, C.c_int, V]
        L.sdpa_gpu_kernel_timing.argtypes = [V, V]
        L.sdpa_gpu_synchronize.argtypes = [V]
        L.sdpa_gpu_mixed_stats.argtypes = [V, np.ctypeslib.ndpointer(np.int64)]
        # lambda* model (include/sdpa_mlp.h)
        L.sdpa_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(V)]
        L.sdpa_mlp_destroy.argtypes = [V]
        L.sdpa_mlp_nparams.restype = C.c_long
        L.sdpa_mlp_nparams.argtypes = [V]
        L.sdpa_mlp_set_params.argtypes = [V, _dp, C.c_long]
        L.sdpa_mlp_get_params.argtypes = [V, _dp, C.c_long]
        L.sdpa_mlp_set_data.argtypes = [V, _dp, _dp, C.c_int]
        L.sdpa_mlp_loss_grad.argtypes = [V, C.POINTER(C.c_double), V]
        L.sdpa_mlp_train.argtypes = [V, C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.sdpa_mlp_predict.argtypes = [V, _dp, C.c_int, _dp]
        L.sdpa_solver_new.restype = V
        L.sdpa_solver_new.argtypes = [C.c_int]
        L.sdpa_solver_free.argtypes = [V]
        L.sdpa_solver_error.restype = C.c_char_p
        L.sdpa_solver_error.argtypes = [V]
        L.sdpa_solver_set_param.argtypes = [V, C.c_char_p, C.c_double]
        L.sdpa_solver_load.argtypes = [V, C.c_int, C.c_int, _ip, _dp, C.c_long, _ip, _ip, _ip, _ip,
                                       _dp]
        L.sdpa_solver_read_input.argtypes = [V, C.c_char_p]
        L.sdpa_solver_read_parameter.argtypes = [V, C.c_char_p]
        L.sdpa_solver_set_init.argtypes = [V, V, V, V]
        L.sdpa_solver_enable_timing.argtypes = [V, C.c_int]
        L.sdpa_solver_solve.argtypes = [V]
        L.sdpa_solver_prepare.argtypes = [V]
        L.sdpa_solver_run.argtypes = [V]
        L.sdpa_solver_result.argtypes = [V, _dp]
        L.sdpa_solver_trace.argtypes = [V, _dp, C.c_int]
        L.sdpa_solver_iterate.argtypes = [V, V, V, V]
        L.sdpa_solver_phase_times.argtypes = [V, _dp]
        L.sdpa_solver_context.restype = V
        L.sdpa_solver_context.argtypes = [V]
        L.sdpa_solver_read_init.argtypes = [V, C.c_char_p]
        L.sdpa_solver_write_input.argtypes = [V, C.c_char_p, C.c_char_p]
        L.sdpa_solver_write_init.argtypes = [V, C.c_char_p, C.c_char_p]
        L.sdpa_solver_dimacs.argtypes = [V, V]
        L.sdpa_solver_print_result.argtypes = [V, C.c_int, C.c_char_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _records(d):
    return (np.ascontiguousarray(d.block_struct, np.int32), np.ascontiguousarray(d.c, np.float64),
            np.ascontiguousarray(d.k, np.int32), np.ascontiguousarray(d.l, np.int32),
            np.ascontiguousarray(d.i, np.int32), np.ascontiguousarray(d.j, np.int32),
            np.ascontiguousarray(d.v, np.float64))


class GpuError(RuntimeError):
    pass


class GpuContext:
    """The kernel-level C ABI (include/sdpa_gpu.h) on one device."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.sdpa_gpu_create(device, C.byref(h))
        if rc != 0:
            raise GpuError(f"sdpa_gpu_create failed ({rc})")
        self.h = h
        self.m = 0
        self.total = 0

    def close(self):
        if getattr(self, "h", None):
            lib().sdpa_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise GpuError(f"{what} failed ({rc}): {lib().sdpa_gpu_last_error(self.h).decode()}")
        return rc

    def load(self, d, kappa: float = 12.0):
        self._keep = _records(d)
        bs, c, k, l, i, j, v = self._keep
        self._chk(lib().sdpa_gpu_load_problem(self.h, d.m, len(bs), bs, c, len(k), k, l, i, j, v,
                                              kappa), "load_problem")
        info = np.zeros(16, np.int64)
        lib().sdpa_gpu_problem_info(self.h, info, 16)
        self.info = info
        self.m = int(info[0])
        self.total = int(info[5])
        self.n_dim = int(info[4])
        return self

    def formula_counts(self):
        return self.info[6:9].copy()

    def get(self, name):
        n = self.m * self.m if name == "B" else (self.m if name in ("y", "dy", "g", "pvec") else self.total)
        out = np.zeros(n)
        self._chk(lib().sdpa_gpu_get_buffer(self.h, BUF[name], out, n), "get_buffer")
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, np.float64).reshape(-1)
        self._chk(lib().sdpa_gpu_set_buffer(self.h, BUF[name], a, a.shape[0]), "set_buffer")

    def bmat(self):
        return self.get("B").reshape(self.m, self.m, order="F")

    def identity_start(self, lam):
        self._chk(lib().sdpa_gpu_identity_start(self.h, lam), "identity_start")

    def inv_chol(self):
        return self._chk(lib().sdpa_gpu_inv_chol(self.h), "inv_chol")

    def make_bmat(self, precision=0):
        return self._chk(lib().sdpa_gpu_make_bmat(self.h, precision), "make_bmat")

    def factor_bmat(self, precision=0):
        return self._chk(lib().sdpa_gpu_factor_bmat(self.h, precision), "factor_bmat")

    def solve_bmat(self, refine=0):
        return self._chk(lib().sdpa_gpu_solve_bmat(self.h, refine), "solve_bmat")

    def min_eig(self):
        a, b = C.c_double(), C.c_double()
        self._chk(lib().sdpa_gpu_min_eig(self.h, C.byref(a), C.byref(b)), "min_eig")
        return a.value, b.value

    def residuals(self):
        a, b = C.c_double(), C.c_double()
        self._chk(lib().sdpa_gpu_residuals(self.h, C.byref(a), C.byref(b)), "residuals")
        return a.value, b.value

    def make_gvec(self, corrector, target, infeasible):
        return self._chk(lib().sdpa_gpu_make_gvec(self.h, int(corrector), float(target),
                                                  int(infeasible)), "make_gvec")

    def make_direction(self, infeasible):
        return self._chk(lib().sdpa_gpu_make_direction(self.h, int(infeasible)), "make_direction")

    def dots(self, *names):
        codes = np.array([DOT[n] for n in names], np.int32)
        out = np.zeros(len(codes))
        self._chk(lib().sdpa_gpu_dots(self.h, len(codes), codes, out), "dots")
        return out

    def update(self, ap, ad):
        return self._chk(lib().sdpa_gpu_update(self.h, ap, ad), "update")

    def min_eig_dots(self, *names):
        """min_eig() and dots(*names) with one host round trip:
        (lamX, lamZ, products)."""
        codes = np.array([DOT[n] for n in names], np.int32)
        out = np.zeros(max(1, len(codes)))
        a, b = C.c_double(), C.c_double()
        self._chk(lib().sdpa_gpu_min_eig_dots(self.h, C.byref(a), C.byref(b), len(codes), codes,
                                              out), "min_eig_dots")
        return a.value, b.value, out[: len(codes)]

    def update_end(self, *names):
        """inv_chol(), dots(*names) and residuals() with one host round trip:
        (status, products, normP, normD); status 1 if an X/Z factorization
        failed (the other outputs are then not set)."""
        codes = np.array([DOT[n] for n in names], np.int32)
        out = np.zeros(max(1, len(codes)))
        a, b = C.c_double(), C.c_double()
        r = self._chk(lib().sdpa_gpu_update_end(self.h, len(codes), codes, out, C.byref(a),
                                                C.byref(b)), "update_end")
        return r, out[: len(codes)], a.value, b.value

    def timing(self, enable=-1):
        out = np.zeros(10)
        lib().sdpa_gpu_timing(self.h, enable, out.ctypes.data_as(C.c_void_p))
        return out

    def synchronize(self):
        self._chk(lib().sdpa_gpu_synchronize(self.h), "synchronize")

    def mixed_stats(self):
        """Mixed-precision counters: factorizations, fallbacks, refined solves,
        refinement steps, stalls, current factor mixed (0/1)."""
        out = np.zeros(6, np.int64)
        self._chk(lib().sdpa_gpu_mixed_stats(self.h, out), "mixed_stats")
        return out


class Result:
    def __init__(self, vals, trace):
        self.primal_obj, self.dual_obj, self.primal_error, self.dual_error, self.mu, \
            self.relative_gap = vals[:6]
        self.iterations = int(vals[6])
        self.phase = PHASES[int(vals[7])]
        self.failure = int(vals[8])
        self.solve_seconds = float(vals[9])
        self.trace = trace

    def user_trace(self):
        """printOneIteration columns with REVERSE_PRIMAL_DUAL (sdpa_io.cpp:638-644)."""
        return [(t[0], t[2], t[1], -t[4], -t[3], t[6], t[5], t[7]) for t in self.trace]


class Solver:
    """Host driver (class sdpa_amd::SDPA) through its C wrapper."""

    def __init__(self, device: int = 0):
        self.h = lib().sdpa_solver_new(device)
        self.m = 0

    def close(self):
        if getattr(self, "h", None):
            lib().sdpa_solver_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise GpuError(f"{what} failed ({rc}): {lib().sdpa_solver_error(self.h).decode()}")
        return rc

    def set_param(self, **kw):
        for k, v in kw.items():
            self._chk(lib().sdpa_solver_set_param(self.h, k.encode(), float(v)), "set_param")
        return self

    def load(self, d):
        self._keep = _records(d)
        bs, c, k, l, i, j, v = self._keep
        self._chk(lib().sdpa_solver_load(self.h, d.m, len(bs), bs, c, len(k), k, l, i, j, v), "load")
        self.m = d.m
        self.total = d.packed_size()
        return self

    def read_input(self, path):
        self._chk(lib().sdpa_solver_read_input(self.h, path.encode()), "read_input")
        return self

    def read_init(self, path):
        """SDPA::readInit (.ini or .ini-s by extension) and setInitPoint(true)."""
        self._chk(lib().sdpa_solver_read_init(self.h, str(path).encode()), "read_init")
        return self

    def write_input(self, path, fmt="%+8.3e"):
        self._chk(lib().sdpa_solver_write_input(self.h, str(path).encode(), fmt.encode()),
                  "write_input")

    def write_init(self, path, fmt="%+8.3e"):
        self._chk(lib().sdpa_solver_write_init(self.h, str(path).encode(), fmt.encode()),
                  "write_init")

    def dimacs(self):
        """DIMACS errors err1..err6 of the final point (index 0 unused)."""
        out = np.zeros(7)
        self._chk(lib().sdpa_solver_dimacs(self.h, out.ctypes.data_as(C.c_void_p)), "dimacs")
        return out

    def print_result(self, which, path):
        """printResultXVec / XMat / YMat (which 0, 1, 2) into a file."""
        self._chk(lib().sdpa_solver_print_result(self.h, int(which), str(path).encode()),
                  "print_result")

    def read_parameter(self, path):
        self._chk(lib().sdpa_solver_read_parameter(self.h, path.encode()), "read_parameter")
        return self

    def set_init(self, X, Z, y=None):
        self._init = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (X, Z, y)]
        self._chk(lib().sdpa_solver_set_init(self.h, *[_ptr(a) for a in self._init]), "set_init")
        return self

    def enable_timing(self, on=True):
        self._chk(lib().sdpa_solver_enable_timing(self.h, 1 if on else 0), "enable_timing")
        return self

    def prepare(self):
        """Upload the problem (SDPA::initializeSolve); later run() calls reuse it."""
        self._chk(lib().sdpa_solver_prepare(self.h), "prepare")
        return self

    def run(self) -> Result:
        """SDPA::solve on the prepared, device-resident problem."""
        self._chk(lib().sdpa_solver_run(self.h), "run")
        return self._result()

    def solve(self) -> Result:
        self._chk(lib().sdpa_solver_solve(self.h), "solve")
        return self._result()

    def _result(self) -> Result:
        vals = np.zeros(12)
        self._chk(lib().sdpa_solver_result(self.h, vals), "result")
        cap = int(vals[6]) + 2
        tr = np.zeros(8 * cap)
        n = lib().sdpa_solver_trace(self.h, tr, cap)
        return Result(vals, tr[: 8 * n].reshape(n, 8))

    def iterate(self, total, m):
        X = np.zeros(total); Z = np.zeros(total); y = np.zeros(m)
        self._chk(lib().sdpa_solver_iterate(self.h, _ptr(X), _ptr(Z), _ptr(y)), "iterate")
        return X, Z, y

    def phase_times(self):
        out = np.zeros(5)
        lib().sdpa_solver_phase_times(self.h, out)
        return dict(zip(("bmat", "chol", "solve", "eig", "inv"), out))

    def kernel_timing(self):
        """HIP-event timing of the Schur Cholesky's two main kernels over the
        last timed solve: {"trailing": (ms total, launches, flops),
        "panel": (ms total, launches, flops)}."""
        ctx = lib().sdpa_solver_context(self.h)
        out = np.zeros(8)
        if ctx:
            lib().sdpa_gpu_kernel_timing(ctx, out.ctypes.data_as(C.c_void_p))
        return {"trailing": (float(out[0]), int(out[1]), float(out[2])),
                "panel": (float(out[4]), int(out[5]), float(out[6]))}

    def mixed_stats(self):
        """Mixed-precision counters of this solver's context (cumulative):
        factorizations, fallbacks to FP64, refined solves, refinement steps,
        stalls, current factor mixed."""
        ctx = lib().sdpa_solver_context(self.h)
        out = np.zeros(6, np.int64)
        if ctx:
            lib().sdpa_gpu_mixed_stats(ctx, out)
        return dict(zip(("factors", "fallbacks", "solves", "refine_steps", "stalls", "current"),
                        (int(v) for v in out)))
