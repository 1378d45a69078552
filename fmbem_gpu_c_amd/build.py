"""Build libfmbem.so in-tree: HIP kernels with hipcc for gfx950, the C host
code with gcc, linked by hipcc into fmbem_gpu_c_amd/libfmbem.so.

Usage: python -m fmbem_gpu_c_amd.build            (or build() from Python)
       python -m fmbem_gpu_c_amd.build --sanitize (libfmbem_san.so: the C host
           code built by gcc with AddressSanitizer and UndefinedBehaviorSanitizer,
           the device code as usual; for CPU runs with
           LD_PRELOAD=$(gcc -print-file-name=libasan.so) and
           FMBEM_LIB=<path of libfmbem_san.so>)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfmbem.so")
OUT_SAN = os.path.join(HERE, "_build_san")
LIB_SAN = os.path.join(OUT_SAN, "libfmbem_san.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer",
             "-fno-sanitize-recover=undefined", "-g"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 ops stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function",
           "-I/opt/rocm/include"]


def _sources():
    hip = sorted(os.path.join(CSRC, "hip", f) for f in os.listdir(os.path.join(CSRC, "hip"))
                 if f.endswith(".hip"))
    c = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
               if f.endswith(".c"))
    return hip, c


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def build(verbose=False, jobs=None, sanitize=False):
    os.makedirs(OUT, exist_ok=True)
    hip, c = _sources()
    cmds, objs = [], []
    for s in hip:
        o = os.path.join(OUT, os.path.basename(s) + ".o")
        if not (sanitize and os.path.exists(o) and os.path.getmtime(o) >= os.path.getmtime(s)):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
        objs.append(o)
    if sanitize:
        os.makedirs(OUT_SAN, exist_ok=True)
    for s in c:
        o = os.path.join(OUT_SAN if sanitize else OUT, os.path.basename(s) + ".o")
        cmds.append([CC] + C_FLAGS + (SAN_FLAGS if sanitize else []) + ["-c", s, "-o", o])
        objs.append(o)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    lib = LIB_SAN if sanitize else LIB
    tmp = lib + ".tmp"
    # the sanitizer runtimes come from LD_PRELOAD (ASan) and libubsan
    extra = ["-lubsan"] if sanitize else []
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + extra)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, sanitize="--sanitize" in sys.argv))
