"""ctypes binding of the C restatement oracle/gather_matvec.c (built into oracle/_build/).

TEST INFRASTRUCTURE ONLY: used by the CPU tests (checked against gfem_oracle.gather_matvec) and
by the CPU-baseline leg of bench.py. gfem/model.py:241-267 restated; see the C source."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "gather_matvec.c")
LIB = os.path.join(HERE, "_build", "liboracle_cpu.so")
_lib = None


def build():
    """gcc -O3 -fopenmp -shared (no -march=native: the library also runs on another host)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    subprocess.run(["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", SRC, "-o", LIB], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.oracle_gather_matvec.restype = None
        lib.oracle_gather_matvec.argtypes = [C.c_int64, C.c_int, C.c_int64] + [C.c_void_p] * 7 + [C.c_int]
        _lib = lib
    return _lib


class GatherMatvec:
    """y = gather_matvec(plan, [(dofs(conn), Ke)], x) for one element block, on `threads` host
    threads. Ke: (ne, nd, nd) FP64 full matrices."""

    def __init__(self, conn, Ke, nn, threads=1):
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.Ke = np.ascontiguousarray(Ke, dtype=np.float64)
        self.ne, self.nnode = self.conn.shape
        self.nn, self.threads = int(nn), int(threads)
        nodes = self.conn.ravel().astype(np.int64)
        self.slot = np.ascontiguousarray(np.argsort(nodes, kind="stable").astype(np.int64))
        self.ptr = np.zeros(self.nn + 1, np.int64)
        self.ptr[1:] = np.cumsum(np.bincount(nodes, minlength=self.nn))
        self.ys = np.empty(self.ne * self.nnode * 3)

    def __call__(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(3 * self.nn)
        load().oracle_gather_matvec(self.ne, self.nnode, self.nn, self.conn.ctypes.data,
                                    self.Ke.ctypes.data, self.ptr.ctypes.data,
                                    self.slot.ctypes.data, x.ctypes.data, self.ys.ctypes.data,
                                    y.ctypes.data, self.threads)
        return y
