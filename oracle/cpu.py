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


def build(path=LIB, native=False):
    """gcc -O3 -fopenmp -shared. native=True adds -march=native: only for a library built on the
    host that runs it (the benchmark's CPU leg builds its own copy there)."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared"] + (["-march=native"] if native else [])
    subprocess.run(cmd + [SRC, "-o", path], check=True)
    return path


def load(path=None):
    """The restatement's shared library: the in-tree build (made if missing), or `path`."""
    global _lib
    if path is not None or _lib is None:
        if path is None:
            path = LIB
            if not os.path.exists(path):
                build()
        lib = C.CDLL(path)
        lib.oracle_gather_matvec.restype = None
        lib.oracle_gather_matvec.argtypes = [C.c_int64, C.c_int, C.c_int64] + [C.c_void_p] * 7 + [C.c_int]
        lib.oracle_jacobi_pcg.restype = C.c_int
        lib.oracle_jacobi_pcg.argtypes = ([C.c_int64, C.c_int, C.c_int64] + [C.c_void_p] * 8 +
                                          [C.c_double, C.c_int] + [C.c_void_p] * 3 + [C.c_int])
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

    def jacobi_pcg(self, b, free, tol=1e-8, maxiter=20000):
        """pcg(opA, b, Jacobi) of gfem/linalg.py:21-47 on this operator with the boundary mask
        `free` (gfem/solvers.py:257-267, pp of :432-434) -> (x, iterations, relres)."""
        free = np.ascontiguousarray(free, dtype=np.float64)
        diag = np.zeros(3 * self.nn)
        nd = 3 * self.nnode
        dofs = (self.conn.astype(np.int64)[:, :, None] * 3 + np.arange(3)).reshape(self.ne, nd)
        np.add.at(diag, dofs.ravel(), np.einsum("eii->ei", self.Ke).ravel())
        pinv = np.ascontiguousarray(np.where(free > 0, 1.0 / np.where(np.abs(diag) > 0, diag, 1.0), 1.0))
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(3 * self.nn)
        work = np.empty(5 * 3 * self.nn)
        rel = C.c_double()
        k = load().oracle_jacobi_pcg(self.ne, self.nnode, self.nn, self.conn.ctypes.data,
                                     self.Ke.ctypes.data, self.ptr.ctypes.data,
                                     self.slot.ctypes.data, free.ctypes.data, pinv.ctypes.data,
                                     b.ctypes.data, x.ctypes.data, float(tol), int(maxiter),
                                     work.ctypes.data, self.ys.ctypes.data, C.byref(rel),
                                     self.threads)
        return x, int(k), rel.value
