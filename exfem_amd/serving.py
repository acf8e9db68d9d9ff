"""Continuous-batching solve server on one MI355X.

Each admitted request gets its own device state - mesh operator, element data, boundary mask,
Krylov workspace and streams - that stays resident in HBM until the request completes. The
scheduler loop never blocks on one solve: every active request advances by a batch of PCG
iterations on its own stream (captured iterations replayed as HIP graphs when set-up runs on
one host thread, enqueued one by one when several set-up threads run), and the loop polls
the requests' events, enqueues the next batch for those whose previous batch has finished,
finishes the converged ones and admits waiting requests as slots and HBM budget free up. Work of
different requests therefore overlaps on the device, and requests arriving while others run are
admitted without waiting for the batch to drain.

Solve forms:
* ``fp64`` with ``inner="ke64"``: the plain path - Jacobi-PCG (the reference's jacobi back-end,
  gfem/solvers.py:432-435) on FP64 upper-triangle element matrices;
* ``fp64`` with ``inner="mf64"``: the same on the FP64 matrix-free operator (element data
  recomputed from the geometry; 2.6x the GDOF/s of the stored form at 1M dof);
* ``cheby32`` (``inner="mf32"``): the mixed-precision served form - FP64 PCG on the FP64
  matrix-free operator (FP64 residual, dot products, norms and update) preconditioned by the
  reference's 6-step Chebyshev-Jacobi polynomial (gfem/solvers.py:289-307) run in FP32 on the FP32
  matrix-free operator (EXFEM_PC_CHEBY_MIXED), with its bound from 25 FP64 power iterations
  (gfem/solvers.py:328-342); about 5x fewer iterations than Jacobi, six of every seven products
  in FP32;
* ``amg32`` (``inner="mf32"``): FP64 PCG on the FP64 matrix-free operator preconditioned by the
  SA-AMG V-cycle run in FP32 (EXFEM_PC_AMG_MIXED: level 0 on the FP32 matrix-free operator,
  coarser levels FP32, coarse solve FP64), its hierarchy set up on the device from the
  request's FP64 element matrices (AMG.from_operator), which are dropped once the hierarchy is
  built; requests below ``amg_min_dof`` (600k dof) take ``small_form`` (default the FP64
  Jacobi form, whose set-up is the cheapest: on the 1e4-2e6-dof stream a hierarchy's set-up pays
  only on the largest boxes, where Jacobi needs the most iterations);
* ``mixed``: FP32 inner Jacobi-PCG (inner tolerance 2e-5) inside FP64 iterative refinement
  (gfem/solvers.py:344-384): the FP64 residual b - A x uses the matrix-free FP64 operator, the
  inner operator is the FP32 matrix-free operator (``inner="mf32"``) or FP32 stored matrices
  (``"ke32"``), all PCG dot products and norms are FP64. Each refinement reduces the FP64
  residual by about kappa(K) eps32, so on ill-conditioned (slender) models it needs more
  iterations in total than the FP64 solve.

A request's result does not depend on what else is in flight: its kernels read only its own
buffers and every reduction has a fixed order.
"""
import collections
import ctypes as C
import time

import numpy as np

from . import _lib as L
from .model import as_dev, stream_ptr, torch, vptr
from .workloads import request_problem


class Request:
    def __init__(self, rid, params, arrival=0.0, problem=None):
        self.id, self.params, self.arrival = rid, params, float(arrival)
        self.problem = problem
        self.t_admit = self.t_done = None
        self.u = None
        self.iters = 0
        self.refinements = 0
        self.relres = None
        self.status = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


def estimate_bytes(ne, nn, mode="mixed", inner="mf32"):
    """HBM held by one active request (element data, operator, workspace, vectors; for amg32
    the FP64 element matrices and the device set-up's peak, then the FP64 + FP32 hierarchy)."""
    ndof = 3 * nn
    b = ne * (8 * 4 + 8 * 3 * 8) + nn * 8 * 4 + ndof * 8 * 12
    if inner == "ke64" or mode == "amg32":
        b += ne * 300 * 8
    else:
        b += ne * 300 * 4          # FP32 matrices (kept only for the diagonal if matrix-free)
    if mode == "amg32":
        b += ndof * 1600
    return int(b * 1.1)


class _Slot:
    """Device state and state machine of one admitted request."""

    # most PCG iterations enqueued at once (the batch doubles up to it): short batches keep the
    # hardware queues that set-up work shares with other requests' solves from backing up
    # (512-request stream, served FP64: 27.7 solves/s at 128, 32.1 at 16)
    max_step = 16

    def __init__(self, req, mode, inner, tol, inner_tol, maxiter, step0, amg_min_dof=600000,
                 small_form="fp64", graphs=True):
        t = torch()
        p0 = req.problem if req.problem is not None else request_problem(req.params)
        if mode == "amg32" and p0.ndof < amg_min_dof:
            mode, inner = small_form, ("mf64" if small_form == "fp64" else inner)
        self.req, self.mode, self.inner = req, mode, inner
        self.amg = None
        t_setup = time.perf_counter()
        self.tol, self.inner_tol, self.maxiter = tol, inner_tol, maxiter
        self.stream = t.cuda.Stream()
        p = req.problem if req.problem is not None else request_problem(req.params)
        req.problem = p
        self.nbytes = estimate_bytes(p.ne, p.nn, mode, inner)
        with t.cuda.stream(self.stream):
            m = p.model()
            self.m = m
            op = m.operator()
            self.op = op
            kdt = L.F64 if (inner == "ke64" or mode == "amg32") else L.F32
            K = m.element_stiffness(storage=L.KE_SYM, dtype=kdt)[0]
            op.set_ke(0, K)
            self.dt = L.F64 if mode in ("fp64", "cheby32", "amg32") else L.F32
            self.mf = inner in ("mf64", "mf32")
            op.set_bc(p.free)
            free = as_dev(p.free.astype(np.float64))
            d = op.diagonal()
            if mode == "amg32":
                from .linalg import AMG
                self.amg = AMG.from_operator(op, p.X, fp32=True)
            if self.mf:
                op.clear_ke(0, kdt)                    # matrix-free: only the diagonal is kept
                del K
            self.pinv = t.where(free > 0, 1.0 / t.where(d.abs() > 0, d, t.ones_like(d)), t.ones_like(d))
            self.lmax = 0.0
            if mode == "cheby32":
                from .linalg import BcOperator, lmax_power
                self.lmax = lmax_power(BcOperator(op, matrix_free=True), self.pinv)
            self.b = as_dev(p.free * p.f)
            self.bn = float(t.linalg.norm(self.b)) or 1.0
            self.x = t.zeros_like(self.b)
            self.r = self.b.clone()
            self.d = t.empty(p.ndof, dtype=t.float64 if self.dt == L.F64 else t.float32,
                             device=self.b.device)
        h = C.c_void_p()
        L.check(L.load().exfem_pcg_create(C.byref(h), p.ndof))
        self.w = h
        if not graphs:
            L.check(L.load().exfem_pcg_set_graph(h, 0))
        req.setup_s = time.perf_counter() - t_setup
        req.form = mode
        self.step = step0
        self.sc = 1.0
        self._begin()

    # ---------------------------------------------------------------- solve steps
    def _begin(self):
        t = torch()
        lib = L.load()
        with t.cuda.stream(self.stream):
            if self.mode in ("fp64", "cheby32", "amg32"):
                rhs, tol = self.b, self.tol
            else:
                self.sc = float(t.linalg.norm(self.r)) or 1.0
                rhs, tol = (self.r / self.sc).float(), self.inner_tol
            self._rhs = rhs
            pc = {"cheby32": L.PC_CHEBY_MIXED, "amg32": L.PC_AMG_MIXED}.get(self.mode, L.PC_JACOBI)
            L.check(lib.exfem_pcg_begin(self.w, self.op.h, vptr(rhs), self.dt, pc,
                                        vptr(self.pinv), self.lmax,
                                        None if self.amg is None else self.amg.h, int(self.mf), tol,
                                        self.maxiter, stream_ptr()))
        L.check(lib.exfem_pcg_advance(self.w, self.step))

    def poll(self):
        """Advance the state machine without blocking on the device. Returns True when the
        request has completed."""
        lib = L.load()
        ready, done, it = C.c_int(), C.c_int(), C.c_int()
        rr = C.c_double()
        st = L.check(lib.exfem_pcg_query(self.w, C.byref(ready), C.byref(done), C.byref(it),
                                         C.byref(rr)))
        if not ready.value:
            return False
        if not done.value:
            self.step = min(self.step * 2, self.max_step)
            L.check(lib.exfem_pcg_advance(self.w, self.step))
            return False
        t = torch()
        self.req.iters += it.value
        with t.cuda.stream(self.stream):
            L.check(lib.exfem_pcg_finish(self.w, vptr(self.d), stream_ptr()))
            if self.mode in ("fp64", "cheby32", "amg32"):
                self.x = self.d.clone()
                self.req.relres = rr.value
                self.req.status = st
                self.stream.synchronize()              # the result is complete in HBM
                return True
            # FP64 refinement: x += sc d, r = b - A x (matrix-free FP64 operator)
            self.x = self.x + self.d.double() * self.sc
            self.r = self.b - self.op.matvec(self.x, bc=True, matrix_free=True)
            self.req.refinements += 1
            rel = float(t.linalg.norm(self.r)) / self.bn
        if rel <= self.tol or self.req.refinements >= 12:
            self.stream.synchronize()
            self.req.relres = rel
            self.req.status = 0 if rel <= self.tol else L.NOT_CONVERGED
            return True
        self.step = 8
        self._begin()
        return False

    def release(self):
        if self.w is not None:
            self.stream.synchronize()
            L.load().exfem_pcg_destroy(self.w)
            self.w = None
        self.op = self.m = self.amg = None


class SolveServer:
    """Continuous-batching server. mode "fp64" (inner "ke64" | "mf64"), "cheby32" (inner
    "mf32"), "amg32" (inner "mf32"; requests below amg_min_dof in small_form, "fp64" (the FP64
    Jacobi form, matrix-free) or "cheby32") or "mixed" (inner "mf32" | "ke32")."""

    def __init__(self, mode="fp64", inner="mf64", max_active=6, mem_budget=None, tol=1e-10,
                 inner_tol=2e-5, maxiter=20000, step0=8, keep_solutions=True, amg_min_dof=600000,
                 setup_threads=3, poll_s=20e-6, small_form="fp64", graphs=None):
        self.mode, self.inner = mode, inner
        self.amg_min_dof = amg_min_dof
        self.small_form = small_form
        # PCG iterations as captured HIP graphs; set-up on more than one host thread enqueues
        # them one by one instead (a capture on one set-up thread fails while another set-up
        # thread makes synchronous HIP calls)
        self.graphs = setup_threads <= 1 if graphs is None else graphs
        if self.graphs and setup_threads > 1:
            raise ValueError("captured PCG iterations need setup_threads <= 1")
        self.setup_threads = setup_threads
        self.poll_s = poll_s
        self.max_active = max_active
        if mem_budget is None:
            free, _ = torch().cuda.mem_get_info()
            mem_budget = int(0.6 * free)
        self.mem_budget = mem_budget
        self.tol, self.inner_tol, self.maxiter, self.step0 = tol, inner_tol, maxiter, step0
        self.keep = keep_solutions

    def run(self, requests, realtime=True, progress=0, log=print):
        """Serve `requests` (their `arrival` times are seconds after the start). Returns the
        completed requests with timing, iterations and (optionally) host solutions. progress > 0
        logs a line every `progress` completed requests. With setup_threads > 0 an admitted
        request's device set-up (operator, element data, AMG hierarchy) runs on one of that many
        host threads while this loop keeps advancing the active solves."""
        import concurrent.futures as cf
        import sys
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active, done, dead, building = [], [], [], []
        used = 0
        self.max_seen, self.max_used = 0, 0
        pool = cf.ThreadPoolExecutor(self.setup_threads) if self.setup_threads else None
        switch = sys.getswitchinterval()
        sys.setswitchinterval(2e-4)            # the loop and the set-up thread share the GIL finely
        try:
            done, wall = self._loop(pending, waiting, active, done, dead, building, pool, requests,
                                    realtime, progress, log)
        finally:
            sys.setswitchinterval(switch)
            if pool is not None:
                pool.shutdown(wait=True)
        return done, wall

    def _loop(self, pending, waiting, active, done, dead, building, pool, requests, realtime, progress, log):
        used = 0
        t0 = time.perf_counter()
        while pending or waiting or active or building:
            now = time.perf_counter() - t0 if realtime else float("inf")
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting and len(active) + len(building) < self.max_active:
                req = waiting[0]
                p = req.problem if req.problem is not None else request_problem(req.params)
                req.problem = p
                small = self.mode == "amg32" and p.ndof < self.amg_min_dof
                need = estimate_bytes(p.ne, p.nn, self.small_form if small else self.mode,
                                      "mf64" if small and self.small_form == "fp64" else self.inner)
                if used + need > self.mem_budget and dead:
                    # freeing device memory synchronises the device: done only when needed
                    for s in dead:
                        s.release()
                        used -= s.nbytes
                    dead = []
                if (active or building) and used + need > self.mem_budget:
                    break
                waiting.popleft()
                req.t_admit = time.perf_counter() - t0
                args = (req, self.mode, self.inner, self.tol, self.inner_tol, self.maxiter, self.step0,
                        self.amg_min_dof, self.small_form, self.graphs)
                if pool is not None:
                    building.append((pool.submit(_Slot, *args), need))
                    used += need
                else:
                    s = _Slot(*args)
                    used += s.nbytes
                    active.append(s)
                self.max_seen = max(self.max_seen, len(active) + len(building))
                self.max_used = max(self.max_used, used)
            for item in list(building):
                if item[0].done():
                    building.remove(item)
                    s = item[0].result()
                    used += s.nbytes - item[1]
                    active.append(s)
            progressed = False
            for s in list(active):
                if s.poll():
                    s.req.t_done = time.perf_counter() - t0
                    if self.keep:
                        with torch().cuda.stream(s.stream):
                            s.req.u = s.x.detach().cpu().numpy()
                    active.remove(s)
                    dead.append(s)
                    done.append(s.req)
                    progressed = True
                    if progress and len(done) % progress == 0:
                        log(f"[serve]   {len(done)}/{len(requests)} done "
                            f"({time.perf_counter() - t0:.1f}s)")
            if not progressed and (active or building):
                time.sleep(self.poll_s)
            elif not active and not building and not waiting and pending and realtime:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
        wall = time.perf_counter() - t0
        for s in dead:
            s.release()
        return done, wall


def summarize(done, wall):
    lat = np.array([r.latency for r in done])
    t_first = min(r.arrival for r in done)
    t_last = max(r.t_done for r in done)
    return dict(requests=len(done), wall_s=wall, solves_per_s=len(done) / max(t_last - t_first, 1e-9),
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                mean_iters=float(np.mean([r.iters for r in done])),
                total_dof=int(sum(r.problem.ndof for r in done)),
                not_converged=int(sum(1 for r in done if r.status)))


def benchmark(n=48, dof_range=(1e4, 2e5), seed=0, max_active=6, load=0.7, tol=1e-10, log=print,
              forms=("served_fp64_mf", "served_amg32_mf", "served_cheby32_mf"), capacity=True, poisson=True,
              rate=None, progress=0, poisson_n=None):
    """Serving measurement on a request stream.
    Capacity (all requests queued at t = 0): two plain FP64 baselines, one request at a time -
    Jacobi-PCG on stored FP64 element matrices (the reference's jacobi back-end) and on the FP64
    matrix-free operator (so batching is measured on its own) - then the continuous-batching
    servers in `forms`: "served_fp64_mf" (FP64 Jacobi, matrix-free), "served_amg32_mf" (FP64 PCG
    with the FP32 SA-AMG V-cycle, hierarchies set up on the device on a host thread; requests
    below 600k dof take the FP64 Jacobi form), "served_cheby32_mf" (FP64 PCG
    with the FP32 Chebyshev polynomial, the mixed-precision form), "served_mixed_mf32" (the
    reference's FP32-inner refinement loop). Every served result's error is stated against the
    plain FP64 solution. Then Poisson arrivals at `rate` requests/s (default: `load` x the plain
    stored capacity) on the first `poisson_n` requests (default all) for the plain stored server
    and the last served form, for p50 / p99 latency. A request arrives with its model (nodes,
    elements, loads, boundary mask); device set-up and solve are timed. Returns a dict."""
    from .workloads import serving_requests
    params = serving_requests(n, seed=seed, dof_range=dof_range)
    t0 = time.perf_counter()
    problems = [request_problem(pr) for pr in params]
    log(f"[serve] {n} request models built on the host ({time.perf_counter() - t0:.1f}s), "
        f"{sum(q.ndof for q in problems)} dof in total")

    def make(arrivals=None, count=None):
        out = []
        for i, pr in enumerate(params[:count]):
            r = Request(pr["id"], pr, 0.0 if arrivals is None else arrivals[i], problems[i])
            out.append(r)
        return out

    res = {}
    kinds = {"served_fp64_mf": ("fp64", "mf64"), "served_cheby32_mf": ("cheby32", "mf32"),
             "served_amg32_mf": ("amg32", "mf32"), "served_mixed_mf32": ("mixed", "mf32")}
    SolveServer("fp64", "ke64", max_active=1, tol=tol, setup_threads=0).run(make()[:1])      # warm-up
    SolveServer("cheby32", "mf32", max_active=1, tol=tol).run(make()[:1])
    if capacity:
        d64, w64 = SolveServer("fp64", "ke64", max_active=1, tol=tol, setup_threads=0).run(
            make(), progress=progress, log=log)
        res["plain_fp64"] = summarize(d64, w64)
        log(f"[serve] plain FP64 (stored, one at a time): {res['plain_fp64']['solves_per_s']:.2f} solves/s")
        dmf, wmf = SolveServer("fp64", "mf64", max_active=1, tol=tol, setup_threads=0).run(
            make(), progress=progress, log=log)
        res["plain_fp64_mf"] = summarize(dmf, wmf)
        log(f"[serve] plain FP64 (matrix-free, one at a time): {res['plain_fp64_mf']['solves_per_s']:.2f} solves/s")
        u64 = {r.id: r.u for r in d64}

        def err(done):
            return max(float(np.linalg.norm(r.u - u64[r.id]) / np.linalg.norm(u64[r.id])) for r in done)
        res["plain_fp64_mf"]["max_rel_diff_vs_plain_fp64"] = err(dmf)
        for name in forms:
            mode, inner = kinds[name]
            srv = SolveServer(mode, inner, max_active=max_active, tol=tol)
            d, w = srv.run(make(), progress=progress, log=log)
            res[name] = summarize(d, w)
            res[name]["max_active_seen"] = srv.max_seen
            res[name]["max_rel_diff_vs_plain_fp64"] = err(d)
            res[name]["speedup_vs_plain_fp64"] = res[name]["solves_per_s"] / res["plain_fp64"]["solves_per_s"]
            res[name]["speedup_vs_plain_fp64_mf"] = res[name]["solves_per_s"] / res["plain_fp64_mf"]["solves_per_s"]
            log(f"[serve] {name}: {res[name]['solves_per_s']:.2f} solves/s "
                f"({res[name]['speedup_vs_plain_fp64']:.2f}x plain stored, "
                f"{res[name]['speedup_vs_plain_fp64_mf']:.2f}x plain matrix-free), "
                f"max rel diff {res[name]['max_rel_diff_vs_plain_fp64']:.1e}, mean iters {res[name]['mean_iters']:.0f}")
        if rate is None:
            rate = load * res["plain_fp64"]["solves_per_s"]
    if poisson:
        if rate is None:
            raise ValueError("the Poisson phase needs a rate (or the capacity phase)")
        npo = n if poisson_n is None else min(n, poisson_n)
        served = forms[-1] if forms else "served_fp64_mf"
        mode, inner = kinds[served]
        arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, npo))
        dp, w = SolveServer("fp64", "ke64", max_active=1, tol=tol, keep_solutions=False, setup_threads=0).run(
            make(arr, npo), progress=progress, log=log)
        ds, w2 = SolveServer(mode, inner, max_active=max_active, tol=tol, keep_solutions=False).run(
            make(arr, npo), progress=progress, log=log)
        res["poisson"] = dict(rate_per_s=rate, requests=npo, served_form=served,
                              plain_fp64=summarize(dp, w), served=summarize(ds, w2))
        log(f"[serve] Poisson {rate:.2f}/s ({npo} requests): plain p50/p99 "
            f"{res['poisson']['plain_fp64']['p50_latency_s']*1e3:.1f}/"
            f"{res['poisson']['plain_fp64']['p99_latency_s']*1e3:.1f} ms, {served} p50/p99 "
            f"{res['poisson']['served']['p50_latency_s']*1e3:.1f}/"
            f"{res['poisson']['served']['p99_latency_s']*1e3:.1f} ms")
    res["stream"] = dict(n=n, dof_range=list(dof_range), seed=seed, tol=tol, max_active=max_active,
                         load=load)
    return res
