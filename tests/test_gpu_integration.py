"""The ctypes binding a gfem maintainer would add runs as written:
the module text is taken from the document, pointed at the in-tree library, and solves the
1,575-dof cantilever with device memory from torch; the result equals the golden FP64 solution
(1e-8) and the Jacobi-PCG iteration count of the product's own path."""
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stub_source():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(# gfem/mi355x\.py\n.*?)```", doc, re.S)
    assert len(blocks) == 1
    return blocks[0]


def test_integration_stub_solves_config0(gpu):
    import torch
    from exfem_amd import _lib, workloads
    src = _stub_source().replace('C.CDLL("libexfem.so")', f'C.CDLL({_lib.LIB_PATH!r})')
    ns = {}
    exec(compile(src, "gfem/mi355x.py", "exec"), ns)
    keep = []

    class Dev:
        @staticmethod
        def upload(a):
            t = torch.as_tensor(np.ascontiguousarray(a), device="cuda")
            keep.append(t)
            return t.data_ptr()

        @staticmethod
        def empty(nbytes):
            t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
            keep.append(t)
            return t.data_ptr()

        @staticmethod
        def download(ptr, n):
            torch.cuda.synchronize()
            for t in keep:
                if t.data_ptr() == ptr:
                    return t.view(torch.float64)[:n].cpu().numpy()
            raise KeyError(ptr)

    p = workloads.small_cantilever()
    sysm = ns["DeviceSystem"](p.X, p.conn, p.E, p.nu, p.free, Dev)
    b = torch.as_tensor(p.free * p.f, device="cuda")
    x = torch.empty_like(b)
    k, rr, st = sysm.solve(b.data_ptr(), x.data_ptr(), 1e-10, 20000)
    torch.cuda.synchronize()
    g = np.load(os.path.join(ROOT, "tests", "golden", "c0_cantilever.npz"))
    u = x.cpu().numpy()
    assert st == 0 and rr <= 1e-10
    assert np.linalg.norm(u - g["u"]) <= 1e-8 * np.linalg.norm(g["u"])
    import exfem_amd as X
    m = p.model()
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    s = X.StaticSolver(m, verbose=False, linear_solver="jacobi", cg_tol=1e-10)
    s.solve(X.Step(bcs=[(fix, d, 0.0) for d in range(3)], dloads=[(p.f,)]))
    assert abs(s.lin.last_iters - k) <= 1
