"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function that include/*.h declares (no compute calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        # C declarations: return type, name, '('
        for m in re.finditer(r"^\s*(?:int|void|long|double|const char\*|sdpa_\w+\*)\s+(sdpa_\w+)\s*\(",
                             txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for must in ("sdpa_gpu_make_bmat", "sdpa_gpu_factor_bmat", "sdpa_gpu_solve_bmat",
                 "sdpa_gpu_inv_chol", "sdpa_gpu_min_eig", "sdpa_solver_solve",
                 "sdpa_batch_solve", "sdpa_mlp_train", "sdpa_mlp_predict"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from sdpa_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("libsdpa_amd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
