"""The host C code of libfmbem.so and the oracle, built with AddressSanitizer
and UndefinedBehaviorSanitizer (python -m fmbem_gpu_c_amd.build --sanitize;
make -C oracle san), run the CPU tests that drive them: the OBJ reader with
its malformed and truncated files, the wrapper checks, the ABI error paths
and the oracle's tree, quadrature and special functions.  Any report ends the
subprocess with a non-zero status (-fno-sanitize-recover)."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

SAN_LIB = os.path.join(ROOT, "fmbem_gpu_c_amd", "_build_san", "libfmbem_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "_san", "liboracle_san.so")


def _asan_runtime():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                           text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _stale(lib, src_dir, exts):
    if not os.path.exists(lib):
        return True
    t = os.path.getmtime(lib)
    for d, _, files in os.walk(src_dir):
        for f in files:
            if f.endswith(exts) and os.path.getmtime(os.path.join(d, f)) > t:
                return True
    return False


def test_cpu_suite_under_asan_ubsan():
    rt = _asan_runtime()
    if not rt:
        pytest.skip("no AddressSanitizer runtime for gcc on this host")
    # the sanitizer builds follow the sources (the HIP objects are reused)
    if _stale(SAN_LIB, os.path.join(ROOT, "fmbem_gpu_c_amd", "csrc"), (".c", ".h", ".hip")):
        from fmbem_gpu_c_amd import build as B
        B.build(sanitize=True)
    if _stale(SAN_ORACLE, os.path.join(ROOT, "oracle"), (".c", ".h")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ, LD_PRELOAD=rt, FMBEM_LIB=SAN_LIB, ORACLE_LIB=SAN_ORACLE,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    tests = ["test_mesh.py", "test_wrapper_checks.py", "test_abi.py", "test_oracle_tree.py",
             "test_oracle_quad.py", "test_oracle_special.py"]
    probe = ("import fmbem_gpu_c_amd as F, oracle as O; F.lib(); O.lib(); "
             "m = open('/proc/self/maps').read(); "
             "assert 'libfmbem_san.so' in m and 'liboracle_san.so' in m and 'libasan' in m")
    p = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True,
                       text=True)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"] + [os.path.join("tests", t) for t in tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-4000:]
    assert "passed" in p.stdout
