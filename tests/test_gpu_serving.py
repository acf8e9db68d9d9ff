"""Serving: results alone == in the mix, FP64 accuracy of the mixed form, memory admission."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu


def _reqs(n=5, seed=3):
    from exfem_amd.serving import Request
    from exfem_amd.workloads import serving_requests
    return [Request(p["id"], p) for p in serving_requests(n, seed=seed, dof_range=(1.5e3, 8e3))]


@pytest.mark.parametrize("setup_threads", [1, 3])
def test_result_is_the_same_alone_and_in_the_mix(gpu, setup_threads):
    """Each request's result is bitwise the same served alone and in the mix: with one set-up
    thread (captured PCG iterations) and with three (iterations enqueued one by one)."""
    from exfem_amd.serving import Request, SolveServer
    for mode, inner in (("fp64", "mf64"), ("mixed", "mf32"), ("cheby32", "mf32"), ("amg32", "mf32")):
        kw = dict(amg_min_dof=1000, setup_threads=setup_threads)
        done, _ = SolveServer(mode, inner, max_active=5, tol=1e-10, **kw).run(_reqs())
        for r in done:
            alone, _ = SolveServer(mode, inner, max_active=1, tol=1e-10, **kw).run([Request(r.id, r.params)])
            assert alone[0].iters == r.iters and alone[0].refinements == r.refinements
            assert np.array_equal(alone[0].u, r.u)


@pytest.mark.parametrize("mode,inner", [("mixed", "mf32"), ("mixed", "ke32"), ("fp64", "ke64"),
                                        ("fp64", "mf64"), ("cheby32", "mf32"), ("amg32", "mf32")])
def test_served_solutions_match_the_fp64_direct_solve(gpu, mode, inner):
    """Every served form == the FP64 direct solve (1e-8); amg32 with every request on its
    device-built hierarchy (threshold 1000 dof)."""
    from exfem_amd.serving import SolveServer
    done, _ = SolveServer(mode, inner, max_active=3, tol=1e-11, amg_min_dof=1000).run(_reqs(3, seed=5))
    if mode == "amg32":
        assert all(r.form == "amg32" for r in done)
    for r in done:
        p = r.problem
        ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], p.E, p.nu))]
        free = p.free.astype(float)
        u0 = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), free * p.f)
        assert np.linalg.norm(r.u - u0) <= 1e-8 * np.linalg.norm(u0), (mode, r.id)
        assert r.status == 0


@pytest.mark.parametrize("small_form", ["fp64", "cheby32"])
def test_amg32_requests_below_the_threshold_take_the_small_form(gpu, small_form):
    """amg32 with a threshold inside the stream's sizes: the small requests are served in
    small_form, the large ones on their hierarchies, each == the FP64 direct solve (1e-8) and
    bitwise the same alone and in the mix."""
    from exfem_amd.serving import Request, SolveServer
    reqs = _reqs(6, seed=11)
    from exfem_amd.workloads import request_problem
    sizes = sorted(request_problem(r.params).ndof for r in reqs)
    thr = (sizes[2] + sizes[3]) // 2
    kw = dict(amg_min_dof=thr, small_form=small_form, tol=1e-11)
    done, _ = SolveServer("amg32", "mf32", max_active=4, **kw).run(reqs)
    forms = {r.id: r.form for r in done}
    for r in done:
        p = r.problem
        assert r.form == (small_form if p.ndof < thr else "amg32"), (r.id, p.ndof, r.form)
        ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], p.E, p.nu))]
        free = p.free.astype(float)
        u0 = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), free * p.f)
        assert np.linalg.norm(r.u - u0) <= 1e-8 * np.linalg.norm(u0), (small_form, r.id)
        assert r.status == 0
        alone, _ = SolveServer("amg32", "mf32", max_active=1, **kw).run([Request(r.id, r.params)])
        assert alone[0].form == forms[r.id] and np.array_equal(alone[0].u, r.u)


def test_admission_respects_the_memory_budget(gpu):
    from exfem_amd.serving import SolveServer, estimate_bytes
    reqs = _reqs(6, seed=7)
    from exfem_amd.workloads import request_problem
    sizes = []
    for r in reqs:
        r.problem = request_problem(r.params)
        sizes.append(estimate_bytes(r.problem.ne, r.problem.nn))
    budget = int(2.2 * max(sizes))
    srv = SolveServer("fp64", "mf64", max_active=6, mem_budget=budget, tol=1e-10)
    done, _ = srv.run(reqs)
    assert len(done) == 6
    assert srv.max_used <= budget or srv.max_seen == 1
    assert srv.max_seen >= 1
