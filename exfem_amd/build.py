"""Build libexfem.so (HIP, gfx950) in-tree: exfem_amd/_lib/libexfem.so.

    python -m exfem_amd.build

Every source compiles to its own object in parallel (one hipcc process per file), then one link
step makes the shared library. Objects go to exfem_amd/_lib/obj/ and are rebuilt on every call.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libexfem.so")
SOURCES = ["elements.hip", "ebe.hip", "krylov.hip", "amg.hip", "block.hip", "discrete.hip", "hyper.hip", "stress.hip", "sa_dev.hip",
           "sa_host.cpp", "patch_host.cpp", "mpc_host.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src, obj, extra, verbose):
    cmd = [hipcc(), *FLAGS, *extra, "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libexfem build failed ({os.path.basename(src)}):\n" + r.stderr[-6000:])
    return obj


def build(verbose=False, extra=(), jobs=None):
    """Compile every HIP / C++ source for gfx950 and link one shared library (always
    recompiles)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 1)), 16)
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, s, o, list(extra), verbose) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *extra, "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libexfem link failed:\n" + r.stderr[-6000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
