"""The C ABI: libfmbem.so loads without a GPU and exports every function
declared in include/fmbem.h with C linkage; argument errors come back as
negative codes with a message (no exit())."""
import os
import re
import subprocess

import fmbem_gpu_c_amd as F
from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "fmbem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fmb_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (fmb_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = F.lib()
    for n in names:
        assert hasattr(L, n)


def test_python_shim_binds_declared_symbols():
    assert set(F._SIGS) <= set(_declared())


def test_error_codes_without_gpu():
    L = F.lib()
    import ctypes
    h = ctypes.c_void_p()
    assert L.fmb_mesh_icosphere(-1, 0.1, ctypes.byref(h)) == -1
    assert b"icosphere" in L.fmb_last_error()
    nv, nf = ctypes.c_int(), ctypes.c_int()
    assert L.fmb_obj_count(b"/nonexistent.obj", ctypes.byref(nv), ctypes.byref(nf)) == -2
    assert L.fmb_trunc_num(18.32, 0.05, 1.5, 0.025) == 14
    assert L.fmb_rot_size(94) == 94 * (4 * 94 * 94 - 1) // 3
    assert L.fmb_coax_size(18) == 18 * (2 * 18 * 18 + 3 * 18 + 1) // 6
    p = F.PlanParams()
    L.fmb_plan_params_default(ctypes.byref(p))
    assert p.eps == 0.05 and p.sigma == 1.5 and p.nedge == 24
