"""Benchmark of the EbE stiffness operator and the PCG solves built on it, one MI355X.

Workload: C3D8 200x40x40 cantilever box of 100x20x20 with interior nodes jittered by
+-10% of the element size (seed 0): 320,000 elements, 337,881 nodes, 1,013,643 dof, clamped at
x = 0, traction (0, 0, -1) on x = 100. One step = one FP64 operator application y = K x over
the whole mesh (element pass streaming the packed element matrices + deterministic node-gather
assembly), inputs resident in HBM. `value` is the GDOF/s of the fastest exact FP64 storage
(upper-triangle element matrices); the full-matrix storage is reported beside it.

Also measured (unless --quick): Jacobi-PCG, SA-AMG-PCG (2-sweep Chebyshev, FP64) and the mixed
FP32-inner / FP64-refinement AMG-PCG to relative residual 1e-8, as solves/s on the same system.

Prints one JSON line with `roofline` (element pass: algorithmic bytes / HIP-event duration vs
8 TB/s) and `cpu_baseline` (the oracle's C gather matvec on the host cores, about 10 s sample).

    python bench.py [--steps K] [--warmup W] [--quick] [--no-cpu] [--calib]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


ROOT = os.path.dirname(os.path.abspath(__file__))
HEAD_KERNEL = "exfem::ebe_elem_kernel<double, 8, true, false>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ebe_elem_sym_fp64_c3d8_200x40x40.traffic.json")
# the sources that define the headline kernel and its launch: the kernel template and its launch
# function in ebe.hip (text between the markers), the row-output helper, the shared constants
KERNEL_SOURCES = ("exfem_amd/csrc/rows.h", "exfem_amd/csrc/common.h")
KERNEL_SLICES = (("exfem_amd/csrc/ebe.hip", "template <typename T, int NN, bool SYM, bool BC>",
                  "// Sign of hexahedral node"),
                 ("exfem_amd/csrc/ebe.hip", "static void launch_elem(", "template <typename T, int ET>"))


def sources_hash():
    """Hash of the headline kernel's own source text: an edit elsewhere in its file leaves the
    committed counter figures valid, an edit of the kernel, its launch or its helpers does not."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update(open(os.path.join(ROOT, f), "rb").read())
    for f, a, b in KERNEL_SLICES:
        t = open(os.path.join(ROOT, f)).read()
        i = t.index(a)
        h.update(t[i:t.index(b, i)].encode())
    return h.hexdigest()


def traffic_from(fetch_csv, write_csv):
    """HBM bytes per launch of the headline kernel from two rocprofv3 counter runs of
    `bench.py --quick --calib` (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by
    what the counter reads for the 256 MiB copy of --calib; written to the traffic file with
    the hash of the kernel sources. Runs on the host, no GPU."""
    import csv
    copy = 256 * 1024        # KiB the calibration copy reads and writes

    def read(path, counter):
        per = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                per.setdefault((r["Kernel_Name"], int(r["Grid_Size"])), []).append(float(r["Counter_Value"]))
        head = [v for (k, _), v in per.items() if HEAD_KERNEL.replace("exfem::", "") in k]
        cal = [v for (k, g), v in per.items() if "copyBuffer" in k and g == 131072]
        if not head or not cal:
            raise SystemExit(f"{path}: no {counter} rows for the headline kernel or the calibration copy")
        return float(np.mean(head[0])), float(np.mean(cal[0]))

    fk, fc = read(fetch_csv, "FETCH_SIZE")
    wk, wc = read(write_csv, "WRITE_SIZE")
    ff, wf = copy / fc, copy / wc
    rec = dict(kernel=HEAD_KERNEL, fetch_kib=fk, write_kib=wk, fetch_factor=ff, write_factor=wf,
               traffic_bytes=(fk * ff + wk * wf) * 1024, sources_sha256=sources_hash(),
               source="profiles/ebe_elem_sym_fp64_c3d8_200x40x40.md")
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec, indent=1))


def time_matvec(op, x, y, steps, warmup, torch, **kw):
    for _ in range(warmup):
        op.matvec(x, y, **kw)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        op.matvec(x, y, **kw)
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def profile_passes(op, x, y, dtype, nrep, mf=False):
    from exfem_amd import _lib as L
    from exfem_amd.model import stream_ptr, vptr
    te, tn = C.c_float(), C.c_float()
    L.check(L.load().exfem_op_profile(op.h, vptr(x), vptr(y), dtype, 0, int(mf), nrep,
                                      C.byref(te), C.byref(tn), stream_ptr()))
    return te.value, tn.value


def mixed_solve(b, A32, M32, A64, tol, inner_tol=2e-5, maxiter=5000, max_ref=12):
    """FP32 AMG-PCG inner solves (tol inner_tol) inside FP64 iterative refinement: residual
    r = b - A64 x and the update in FP64, at most max_ref steps, stop at |r| <= tol |b|
    (gfem/solvers.py:344-384). Returns (x, inner iterations, refinements, relres)."""
    import torch
    from exfem_amd.linalg import pcg_device
    xm = torch.zeros_like(b)
    r = b.clone()
    bn = float(torch.linalg.norm(b))
    its = steps = 0
    for _ in range(max_ref):
        sc = float(torch.linalg.norm(r)) or 1.0
        dd, kk, _, _ = pcg_device(A32, (r / sc).float(), M32, tol=inner_tol, maxiter=maxiter)
        its += kk
        steps += 1
        xm = xm + dd.double() * sc
        r = b - A64(xm)
        if float(torch.linalg.norm(r)) <= tol * bn:
            break
    return xm, its, steps, float(torch.linalg.norm(r)) / bn


def tet10_workload(args, log, lite=False):
    """C3D10 Delaunay box (lattice nx x ny x nz over 1 x 1 x 4, jittered), per-element E
    log-uniform in [1e3, 1e6]: matvecs, AMG setup, FP64 AMG-PCG (tol 1e-12, the FP64 reference
    solution) and the mixed FP32-inner / FP64-refinement AMG-PCG (tol 1e-10), with the relative
    displacement error of the mixed solution."""
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, pcg_device
    t0 = time.time()
    p = workloads.tet10_box(*args.lattice, seed=0)
    t_mesh = time.time() - t0
    log(f"[tet10] mesh {p.ne} C3D10, {p.ndof} dof ({t_mesh:.1f}s)")
    m = X.Model(p.X)
    m.add_block(p.conn, X.TET10, X.LinearElastic(1.0, p.nu))
    fields = dict(scale=[p.scale])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    K64 = m.element_stiffness(fields=fields, storage=L.KE_SYM, dtype=L.F64)[0]
    torch.cuda.synchronize()
    t_ke = time.perf_counter() - t0
    K32 = m.element_stiffness(fields=fields, storage=L.KE_SYM, dtype=L.F32)[0]
    op = m.operator(fields)
    op.set_ke(0, K64)
    op.set_ke(0, K32)
    dev = torch.device("cuda")
    x = torch.as_tensor(np.random.default_rng(1).standard_normal(p.ndof), device=dev)
    y = torch.empty_like(x)
    out = dict(workload=f"C3D10 Delaunay {args.lattice}, E log-uniform [1e3,1e6]", ne=p.ne,
               ndof=p.ndof, mesh_s=t_mesh, ke_compute_s=t_ke)
    mv = {}
    for name, kw in (("sym_f64", dict()), ("sym_f32", dict(dtype=L.F32)),
                     ("mf_f64", dict(matrix_free=True)), ("mf_f32", dict(matrix_free=True, dtype=L.F32))):
        xx = x if kw.get("dtype", L.F64) == L.F64 else x.float()
        yy = torch.empty_like(xx)
        ms = time_matvec(op, xx, yy, 10, 2, torch, **kw)
        mv[name] = dict(us=ms * 1e3, gdofs=p.ndof / (ms * 1e-3) / 1e9)
        log(f"[tet10] matvec {name}: {ms * 1e3:.1f} us ({mv[name]['gdofs']:.2f} GDOF/s)")
    out["matvec"] = mv
    op.set_bc(p.free)
    free = torch.as_tensor(p.free.astype(np.float64), device=dev)
    d = op.diagonal()
    pinv = torch.where(free > 0, 1.0 / torch.where(d.abs() > 0, d, torch.ones_like(d)), torch.ones_like(d))
    b = torch.as_tensor(p.free * p.f, device=dev)
    torch.cuda.synchronize()
    amg = AMG.from_operator(op, p.X, log=log)      # assembly with the mask and SA on the device
    torch.cuda.synchronize()
    out["amg_setup"] = dict(assemble_s=amg.setup_parts["assemble_s"],
                            hierarchy_s=amg.setup_time - amg.setup_parts["assemble_s"],
                            parts=amg.setup_parts, levels=amg.sizes)
    log(f"[tet10] AMG setup on the device: assemble {amg.setup_parts['assemble_s']:.2f}s, hierarchy "
        f"{amg.setup_time - amg.setup_parts['assemble_s']:.2f}s {amg.sizes} "
        + ", ".join(f"{k} {v:.2f}" for k, v in amg.setup_parts.items()))
    A64 = BcOperator(op)
    x64 = None
    if not lite:
        M64 = AMGPrecond(amg, A64, pinv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x64, k64, r64, st = pcg_device(A64, b, M64, tol=1e-12, maxiter=5000)
        torch.cuda.synchronize()
        out["amg_fp64_tol1e-12"] = dict(iters=k64, relres=r64, seconds=time.perf_counter() - t0)
        log(f"[tet10] FP64 AMG-PCG {k64} it relres {r64:.1e} {out['amg_fp64_tol1e-12']['seconds']:.2f}s")
    # level 0 (and every outer product) matrix-free: the same solve without stored matrices
    A64f = BcOperator(op, matrix_free=True)
    M64f = AMGPrecond(amg, A64f, pinv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64f, k64f, r64f, _ = pcg_device(A64f, b, M64f, tol=1e-12, maxiter=5000)
    torch.cuda.synchronize()
    out["amg_fp64_mf_tol1e-12"] = dict(iters=k64f, relres=r64f, seconds=time.perf_counter() - t0)
    if x64 is not None:
        out["amg_fp64_mf_tol1e-12"]["rel_diff_vs_stored"] = float(torch.linalg.norm(x64f - x64) / torch.linalg.norm(x64))
    else:
        x64 = x64f                                 # the FP64 reference solution
    log(f"[tet10] FP64 AMG-PCG matrix-free {k64f} it relres {r64f:.1e} "
        f"{out['amg_fp64_mf_tol1e-12']['seconds']:.2f}s")
    for name, mf in ((("mixed_mf", True),) if lite else (("mixed", False), ("mixed_mf", True))):
        A32 = BcOperator(op, matrix_free=mf, dtype=L.F32)
        M32 = AMGPrecond(amg, A32, pinv, L.F32)
        A64r = A64f if mf else A64
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xm, its, nref, rel = mixed_solve(b, A32, M32, A64r, 1e-10)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        err = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64))
        out[name] = dict(inner_iters=its, refinements=nref, seconds=tm, relres=rel,
                         rel_disp_err_vs_fp64=err)
        log(f"[tet10] {name}: {its} inner it, {nref} refinements, {tm:.2f}s, rel disp err {err:.1e}")
    # FP64 PCG preconditioned by the FP32 V-cycle, tol 1e-10
    for name, A in ((("fp64_cycle32_mf", A64f),) if lite else (("fp64_cycle32", A64), ("fp64_cycle32_mf", A64f))):
        Mc = AMGPrecond(amg, A, pinv, L.F64, cycle=L.F32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xc, kc, rc, _ = pcg_device(A, b, Mc, tol=1e-10, maxiter=5000)
        torch.cuda.synchronize()
        tc = time.perf_counter() - t0
        err = float(torch.linalg.norm(xc - x64) / torch.linalg.norm(x64))
        out[name] = dict(iters=kc, relres=rc, seconds=tc, rel_disp_err_vs_fp64=err)
        log(f"[tet10] {name}: {kc} it, relres {rc:.1e}, {tc:.2f}s, rel disp err {err:.1e}")
    return out


def modal_workload(args, log):
    """Natural frequencies of the jittered C3D8 box (clamped at x = 0, E 210000, nu 0.3, steel
    density 7.85e-9): the device LOBPCG (batched EbE product, SA-AMG V-cycle preconditioner)."""
    import torch
    import exfem_amd as X
    from exfem_amd import workloads
    nx, ny, nz = args.nx
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=0)
    m = X.Model(p.X)
    m.add_block(p.conn, X.HEX8, X.LinearElastic(210000.0, 0.3, density=7.85e-9))
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    st = X.Step(kind="frequency", bcs=[(fix, d, 0.0) for d in range(3)], nmodes=args.nmodes,
                eigensolver="lobpcg")
    mon = X.Monitor(False)
    mon.log = log
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = X.modal(m, st, monitor=mon)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"[modal] {p.ndof} dof, {args.nmodes} modes: {r['iterations']} LOBPCG iterations, {dt:.2f}s")
    return dict(workload=f"modal, C3D8 {nx}x{ny}x{nz} jittered box", ndof=p.ndof, nmodes=args.nmodes,
                seconds=dt, iterations=r["iterations"], freq_hz=[float(f) for f in r["freq"]])


def j2_newton_workload(args, log):
    """The Newton path at scale: the jittered C3D8 cantilever (100 x 20 x 20, clamped at x = 0)
    with J2Plasticity(210e3, 0.3, [[250, 0], [400, 0.2]]) and the tip traction (0, 0, -t), t
    about 1.45x first yield at the root (6 t L / h = 250 at t = 8.3), in 4 equal increments:
    StaticSolver with the consistent tangent re-formed at every iteration and AMG-PCG
    (tol 1e-8) for every linear solve."""
    import torch
    import exfem_amd as X
    from exfem_amd import workloads
    nx, ny, nz = args.nx
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=0)
    m = X.Model(p.X)
    m.add_block(p.conn, X.HEX8, X.J2Plasticity(210e3, 0.3, [[250, 0], [400, 0.2]]))
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    t = args.j2_traction
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], dloads=[(t * p.f,)], dt_init=0.25,
                dt_max=0.25)
    mon = X.Monitor(False)
    mon.log = log
    s = X.StaticSolver(m, monitor=mon, linear_solver="amg", cg_tol=1e-8, tangent=args.tangent)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, states, info = s.solve(st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    eq = states[0][..., 9]
    tip = float(u[2::3].min())
    iters = [it for _, it in info["history"]]
    log(f"[j2newton] {p.ndof} dof, tangent {args.tangent}: {len(iters)} increments, Newton iterations {iters}, "
        f"{dt:.2f}s; tip w {tip:.4f}, plastic points {float((eq > 0).float().mean()):.1%}, "
        f"max eps_p {float(eq.max()):.2e}")
    return dict(workload=f"J2 cantilever C3D8 {nx}x{ny}x{nz} jittered, traction {t}", ndof=p.ndof,
                tangent=args.tangent,
                seconds=dt, newton_iterations=iters, tip_w=tip,
                plastic_fraction=float((eq > 0).float().mean()), max_eqps=float(eq.max()))


def block_workload(args, log):
    """Config 3 alone: the 64 load cases (top-face N(0, 1) tractions, default_rng(k)) on the
    jittered C3D8 box, block PCG at tol 1e-8 with --precond (jacobi: the reference's; amg / amg32:
    the batched SA-AMG V-cycle, FP64 or FP32 cycle). The operator, element matrices and AMG
    hierarchy are built once; each of --reps solves is timed whole."""
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L, workloads
    from exfem_amd.linalg import AMG
    from exfem_amd.optimize import multi_load_cases
    from exfem_amd.solvers import State, _Setup
    nx, ny, nz = args.nx
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=0)
    m = p.model()
    F = workloads.top_face_loads(p, 64)
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    step = X.Step(bcs=[(fix, d, 0.0) for d in range(3)])
    amg = None
    if args.precond != "jacobi":
        op = m.operator()
        for i, K in enumerate(m.element_stiffness(dtype=L.F64)):
            op.set_ke(i, K)
        op.set_bc(_Setup(m, step, State(m)).free_u8)
        amg = AMG.from_operator(op, m.X)
    out = []
    for rep in range(args.reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        U, info = multi_load_cases(m, step, F, cg_tol=1e-8, return_info=True, precond=args.precond,
                                   amg=amg)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out.append(dt)
        log(f"[block] {args.precond} rep {rep}: {dt:.3f}s, iters {info['iters'].min()}-{info['iters'].max()}")
    return dict(workload=f"64 load cases, C3D8 {nx}x{ny}x{nz} jittered box", precond=args.precond,
                seconds=out, max_iters=int(info["iters"].max()))


def j2_workload(args, log):
    """The kernel benchmark of the reference (examples/benchmark.py:21-34) on the device: a
    unit-spaced n^3 C3D8 box with J2Plasticity(210e3, 0.3, [[250, 0], [400, 0.2]]), zero states,
    u ~ 1e-3 N(0, 1) (NumPy default_rng(0)); times the internal force (return at every Gauss
    point, states written), the consistent-tangent element matrices (full FP64 storage, as the
    reference returns them) and the EbE product with them, each the mean of repeats after one
    warm-up, and the matrix-free consistent tangent K_t(u, s) v (the reference's JVP column):
    its set-up at (u, s) - the return at every Gauss point, once per linearisation point - and
    the product. The kernels are timed from HIP events on the stream they run on."""
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L
    from exfem_amd.model import stream_ptr
    n = args.j2_n
    Xn, conn = X.mesh.box_hex(n, n, n, n, n, n)
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.J2Plasticity(210e3, 0.3, [[250, 0], [400, 0.2]]))
    s = m.init_state()
    u = torch.as_tensor(np.random.default_rng(0).standard_normal(m.ndof) * 1e-3, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, rep):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(rep):
            out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / rep, out

    t_f, (f, s1) = timed(lambda: m.internal_force(u, s), 10)
    t_k, Ks = timed(lambda: m.element_stiffness(u, s), 5)
    op = m.operator()
    op.set_ke(0, Ks[0])
    t_v, _ = timed(lambda: op.matvec(u), 20)
    t_js, _ = timed(lambda: op.set_state(u, s), 10)
    t_j, _ = timed(lambda: op.matvec(u, matrix_free=True), 20)
    plastic = float((s1[0][..., 9] > 0).float().mean())
    ne = m.blocks[0].ne
    log(f"[j2] {ne} hex8 {m.ndof} dofs | fint {t_f:.3f} ms | Ke {t_k:.3f} ms | EbE K*v {t_v:.3f} ms "
        f"| JVP set-up {t_js:.3f} ms, K*v {t_j:.3f} ms | plastic points {plastic:.1%}")
    # traffic floors: the force reads connectivity, coordinates, u and the old state and writes
    # the new state (19 f64 per point) and the partial rows; Ke writes 576 f64 per element
    f_bytes = ne * (8 * 4 + 2 * 8 * 19 * 8) + m.nn * 24 * 2 + m.ndof * 8 * 2
    k_bytes = ne * (8 * 4 + 8 * 19 * 8 + 576 * 8) + m.nn * 24 + m.ndof * 8
    return dict(workload=f"J2 C3D8 {n}^3 box (examples/benchmark.py)", ne=ne, ndof=m.ndof,
                fint_ms=t_f, ke_ms=t_k, kv_ms=t_v, jvp_setup_ms=t_js, jvp_ms=t_j,
                plastic_fraction=plastic,
                fint_bytes=f_bytes, fint_hbm_frac=f_bytes / (t_f * 1e-3) / 8e12,
                ke_bytes=k_bytes, ke_hbm_frac=k_bytes / (t_k * 1e-3) / 8e12)


def hyper_workload(args, log):
    """The J2 kernel benchmark's box (unit-spaced n^3 C3D8, u ~ 1e-3 N(0, 1)) with each
    hyperelastic material at E = 210e3, nu = 0.3 (finite strain, F-bar): the internal force, the
    element tangents (material part and the F-bar geometric term, full FP64 storage), the EbE
    product with them and the matrix-free tangent product (the dual-number derivative of the
    force at the point set by set_state), each the mean of repeats after one warm-up, timed from
    HIP events on the stream they run on."""
    import torch
    import exfem_amd as X
    n = args.j2_n
    Xn, conn = X.mesh.box_hex(n, n, n, n, n, n)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, rep):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(rep):
            out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / rep, out

    mats = [("NeoHookean", X.NeoHookean(E=210e3, nu=0.3)),
            ("MooneyRivlin", X.MooneyRivlin(C10=30e3, C01=10e3, D1=2e-5)),
            ("ArrudaBoyce", X.ArrudaBoyce(mu=80e3, lam_m=3.0, D=2e-5))]
    out = {}
    for name, mat in mats:
        m = X.Model(Xn)
        m.add_block(conn, X.HEX8, mat)
        u = torch.as_tensor(np.random.default_rng(0).standard_normal(m.ndof) * 1e-3, device="cuda")
        t_f, _ = timed(lambda: m.internal_force(u), 10)
        t_k, Ks = timed(lambda: m.element_stiffness(u), 3)
        op = m.operator()
        op.set_ke(0, Ks[0])
        t_v, _ = timed(lambda: op.matvec(u), 20)
        op.set_state(u, [None])
        t_j, _ = timed(lambda: op.matvec(u, matrix_free=True), 10)
        ne = m.blocks[0].ne
        log(f"[hyper] {name}: {ne} hex8 {m.ndof} dofs | fint {t_f:.3f} ms | Ke {t_k:.3f} ms | "
            f"EbE K*v {t_v:.3f} ms | matrix-free K*v {t_j:.3f} ms")
        out[name] = dict(ne=ne, ndof=m.ndof, fint_ms=t_f, ke_ms=t_k, kv_ms=t_v, jvp_ms=t_j)
        del op, Ks, m
    out["workload"] = f"hyperelastic C3D8 {n}^3 box, u ~ 1e-3 N(0,1)"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, nargs=3, default=[200, 40, 40])
    ap.add_argument("--quick", action="store_true", help="matvec only (no PCG solves)")
    ap.add_argument("--no-amg", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-multi", action="store_true", help="skip the 64-case measurements")
    ap.add_argument("--no-block", action="store_true",
                    help="64-case section: time the batched product only (no block PCG)")
    ap.add_argument("--j2-n", type=int, default=69, help="J2 workload: elements per box edge")
    ap.add_argument("--precond", choices=["jacobi", "amg", "amg32"], default="amg",
                    help="block workload: preconditioner of the 64-case block PCG")
    ap.add_argument("--reps", type=int, default=2, help="block workload: timed solves")
    ap.add_argument("--j2-traction", type=float, default=12.0, help="J2 Newton workload: tip traction")
    ap.add_argument("--tangent", choices=["ebe", "jvp"], default="ebe",
                    help="J2 Newton workload: element-matrix or matrix-free tangent products")
    ap.add_argument("--workload", choices=["hex", "tet10", "serve", "modal", "j2", "block", "j2newton",
                                           "hyper"],
                    default="hex",
                    help="tet10: the C3D10 log-uniform-E mixed-precision workload instead; "
                         "serve: only the serving stream (--serve-n requests, dof up to "
                         "--serve-dof-max)")
    ap.add_argument("--lattice", type=int, nargs=3, default=[42, 42, 168])
    ap.add_argument("--nmodes", type=int, default=10)
    ap.add_argument("--serve-n", type=int, default=512)
    ap.add_argument("--serve-dof-max", type=float, default=2e6)
    ap.add_argument("--serve-poisson-n", type=int, default=128,
                    help="serving: requests of the Poisson-arrival phase")
    ap.add_argument("--no-tet10", action="store_true", help="skip the 8.1M-dof C3D10 leg")
    ap.add_argument("--serve-phase", choices=["all", "capacity", "poisson"], default="all")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="Poisson arrival rate (requests/s); default 0.7 x the plain capacity")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline sample length")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the headline kernel's HBM traffic from two rocprofv3 counter CSVs")
    ap.add_argument("--calib", action="store_true",
                    help="also run one plain 256 MiB device copy (counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the operator would "
              "shard by element slabs with a halo exchange); run it with --gpus 1", flush=True)
        sys.exit(2)
    if args.traffic_from:
        traffic_from(*args.traffic_from)
        return

    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads

    L.load()
    if args.workload == "tet10":
        print(json.dumps(dict(metric="C3D10 mixed-precision AMG-PCG", extra=tet10_workload(args, log))),
              flush=True)
        return
    if args.workload == "modal":
        print(json.dumps(dict(metric="modal analysis (device LOBPCG)", extra=modal_workload(args, log))),
              flush=True)
        return
    if args.workload == "j2newton":
        print(json.dumps(dict(metric="J2 Newton static solve (s)", extra=j2_newton_workload(args, log))),
              flush=True)
        return
    if args.workload == "block":
        print(json.dumps(dict(metric="64-case block PCG (s)", extra=block_workload(args, log))), flush=True)
        return
    if args.workload == "hyper":
        print(json.dumps(dict(metric="hyperelastic internal force / tangents (ms)",
                              extra=hyper_workload(args, log))), flush=True)
        return
    if args.workload == "j2":
        print(json.dumps(dict(metric="J2 internal force / element tangents (ms)", extra=j2_workload(args, log))),
              flush=True)
        return
    if args.workload == "serve":
        from exfem_amd import serving
        res = serving.benchmark(n=args.serve_n, dof_range=(1e4, args.serve_dof_max), log=log,
                                forms=("served_fp64_mf", "served_amg32_mf", "served_cheby32_mf"),
                                capacity=args.serve_phase != "poisson",
                                poisson=args.serve_phase != "capacity", rate=args.serve_rate,
                                progress=16, poisson_n=args.serve_poisson_n)
        print(json.dumps(dict(metric="served solves/s and p50/p99 latency", extra=res)), flush=True)
        return
    dev = torch.device("cuda")
    nx, ny, nz = args.nx
    t0 = time.time()
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=0)
    log(f"[bench] mesh {p.name}: ne={p.ne} nn={p.nn} ndof={p.ndof} ({time.time() - t0:.1f}s)")
    m = p.model()
    op = m.operator()
    x = torch.as_tensor(np.random.default_rng(1).standard_normal(p.ndof), device=dev)
    y = torch.empty_like(x)
    extra = {}
    nrep = max(10, args.steps // 2)

    # ---- element matrices: full and upper-triangle storage, FP64; FP32 upper triangle
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Kf = m.element_stiffness(storage=L.KE_FULL)[0]
    torch.cuda.synchronize()
    extra["ke_compute_full_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    Ks = m.element_stiffness(storage=L.KE_SYM)[0]
    torch.cuda.synchronize()
    extra["ke_compute_sym_ms"] = (time.perf_counter() - t0) * 1e3
    Ks32 = m.element_stiffness(storage=L.KE_SYM, dtype=L.F32)[0]

    s8, nnode = 8, 8
    res = {}
    for name, K, ss, nent in (("full", Kf, 8, 576), ("sym", Ks, 8, 300), ("sym_f32", Ks32, 4, 300)):
        op.set_ke(0, K)
        dt = L.F64 if ss == 8 else L.F32
        xx = x if dt == L.F64 else x.float()
        yy = y if dt == L.F64 else torch.empty_like(xx)
        ms = time_matvec(op, xx, yy, args.steps, args.warmup, torch, dtype=dt)
        te, tn = profile_passes(op, xx, yy, dt, nrep)
        b_elem = p.ne * (nent * ss + nnode * 4) + p.ndof * ss
        b_mv = p.ne * (nent * ss + nnode * 4) + 2 * p.ndof * ss
        res[name] = dict(us_per_matvec=ms * 1e3, gdofs=p.ndof / (ms * 1e-3) / 1e9,
                         matvec_bytes=b_mv, matvec_hbm_frac=b_mv / (ms * 1e-3) / HBM_PEAK,
                         elem_pass_us=te * 1e3, node_pass_us=tn * 1e3, elem_bytes=b_elem,
                         elem_hbm_frac=b_elem / (te * 1e-3) / HBM_PEAK)
        log(f"[bench] matvec {name}: {ms * 1e3:.1f} us ({res[name]['gdofs']:.2f} GDOF/s, "
            f"elem {te * 1e3:.1f} us = {res[name]['elem_hbm_frac'] * 100:.1f}% HBM, node {tn * 1e3:.1f} us)")
    ms = time_matvec(op, x, y, args.steps, args.warmup, torch, matrix_free=True)
    te, tn = profile_passes(op, x, y, L.F64, nrep, mf=True)
    res["matrix_free"] = dict(us_per_matvec=ms * 1e3, gdofs=p.ndof / (ms * 1e-3) / 1e9,
                              elem_pass_us=te * 1e3, node_pass_us=tn * 1e3)
    log(f"[bench] matvec matrix-free: {ms * 1e3:.1f} us ({res['matrix_free']['gdofs']:.2f} GDOF/s)")
    # the assembled operator (operator="csr"): 3x3 block CSR summed on the device from the FP64 /
    # FP32 upper-triangle element matrices; bytes = blocks (9 values + column) + row offsets + x + y
    op.set_ke(0, Ks)
    op.set_ke(0, Ks32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.set_assembled(True)
    op.matvec(x, y)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    nblk = op.bsr_nblocks()
    asm = dict(assembly_ms=t_asm * 1e3, nblocks=nblk)
    for name, dt, ss in (("fp64", L.F64, 8), ("fp32", L.F32, 4)):
        xx = x if dt == L.F64 else x.float()
        yy = y if dt == L.F64 else torch.empty_like(xx)
        ms = time_matvec(op, xx, yy, args.steps, args.warmup, torch, dtype=dt)
        byt = nblk * (9 * ss + 4) + (p.nn + 1) * 4 + 2 * p.ndof * ss
        asm[name] = dict(us_per_matvec=ms * 1e3, gdofs=p.ndof / (ms * 1e-3) / 1e9, bytes=byt,
                         hbm_frac=byt / (ms * 1e-3) / HBM_PEAK)
        log(f"[bench] matvec assembled {name}: {ms * 1e3:.1f} us ({asm[name]['gdofs']:.2f} GDOF/s, "
            f"{asm[name]['hbm_frac'] * 100:.1f}% HBM on {byt / 1e6:.0f} MB)")
    log(f"[bench] assembled operator: {nblk} blocks, pattern + values {t_asm * 1e3:.0f} ms")
    op.set_assembled(False)
    res["assembled"] = asm
    extra["matvec"] = res
    if args.calib:
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    head = res["sym"]
    # measured HBM bytes per launch of this kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE, each scaled
    # by a plain-copy calibration; --traffic-from), committed beside the kernel-trace summary and
    # used only while the kernel's sources are the ones the counters were taken on
    traffic, note = None, None
    if (nx, ny, nz) == (200, 40, 40) and os.path.exists(TRAFFIC_FILE):
        rec = json.load(open(TRAFFIC_FILE))
        if rec.get("sources_sha256") == sources_hash():
            traffic = rec["traffic_bytes"]
        else:
            note = "counters predate the current kernel sources: re-measure (--traffic-from)"
    roofline = dict(bound="hbm", kernel=HEAD_KERNEL,
                    achieved=head["elem_bytes"] / (head["elem_pass_us"] * 1e-6) / 1e9,
                    peak=HBM_PEAK / 1e9, unit="GB/s", frac=head["elem_hbm_frac"],
                    traffic=traffic, algorithmic_bytes=head["elem_bytes"],
                    elem_pass_us=head["elem_pass_us"])
    if note:
        roofline["traffic_note"] = note

    # ---- solves on the same system (reference stopping rule, tol 1e-8)
    if not args.quick:
        from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, Jacobi, pcg_device
        from exfem_amd.model import ops_assemble
        op.set_ke(0, Ks)
        op.set_ke(0, Ks32)
        op.set_bc(p.free)
        free = torch.as_tensor(p.free.astype(np.float64), device=dev)
        d = op.diagonal()
        pinv = torch.where(free > 0, 1.0 / torch.where(d.abs() > 0, d, torch.ones_like(d)),
                           torch.ones_like(d))
        bvec = torch.as_tensor(p.free * p.f, device=dev)
        A64 = BcOperator(op)
        solves = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xj, k, rr, st = pcg_device(A64, bvec, Jacobi(pinv), tol=1e-8, maxiter=20000)
        torch.cuda.synchronize()
        dtj = time.perf_counter() - t0
        solves["jacobi_fp64"] = dict(iters=k, relres=rr, seconds=dtj, solves_per_s=1 / dtj)
        log(f"[bench] Jacobi-PCG {k} it, relres {rr:.2e}, {dtj:.3f}s")
        if not args.no_amg:
            op.set_ke(0, Ks)
            torch.cuda.synchronize()
            amg = AMG.from_operator(op, p.X)              # assembly and SA set-up on the device
            torch.cuda.synchronize()
            t_asm = amg.setup_parts["assemble_s"]
            solves["amg_setup"] = dict(assemble_s=t_asm, hierarchy_s=amg.setup_time - t_asm,
                                       parts=amg.setup_parts, levels=amg.sizes,
                                       complexity=amg.complexity)
            log(f"[bench] AMG setup on the device: assemble {t_asm:.2f}s, hierarchy "
                f"{amg.setup_time - t_asm:.2f}s {amg.sizes}")
            M64 = AMGPrecond(amg, A64, pinv)
            pcg_device(A64, bvec, M64, tol=1e-8, maxiter=2000)           # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xa, k, rr, st = pcg_device(A64, bvec, M64, tol=1e-8, maxiter=2000)
            torch.cuda.synchronize()
            dta = time.perf_counter() - t0
            solves["amg_fp64"] = dict(iters=k, relres=rr, seconds=dta, solves_per_s=1 / dta)
            log(f"[bench] AMG-PCG FP64 {k} it, relres {rr:.2e}, {dta * 1e3:.1f} ms")
            # level 0 and every outer product matrix-free (no stored matrices in the solve)
            A64f = BcOperator(op, matrix_free=True)
            M64f = AMGPrecond(amg, A64f, pinv)
            pcg_device(A64f, bvec, M64f, tol=1e-8, maxiter=2000)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xf, k, rr, st = pcg_device(A64f, bvec, M64f, tol=1e-8, maxiter=2000)
            torch.cuda.synchronize()
            dtf = time.perf_counter() - t0
            solves["amg_fp64_mf"] = dict(iters=k, relres=rr, seconds=dtf, solves_per_s=1 / dtf,
                                         rel_diff_vs_stored=float(torch.linalg.norm(xf - xa) / torch.linalg.norm(xa)))
            log(f"[bench] AMG-PCG FP64 matrix-free {k} it, relres {rr:.2e}, {dtf * 1e3:.1f} ms")
            # mixed: FP32 AMG-PCG inner (tol 2e-5), FP64 residual / refinement (<= 12 steps)
            for name, mf in (("mixed_fp32_fp64", False), ("mixed_fp32_fp64_mf", True)):
                A32 = BcOperator(op, matrix_free=mf, dtype=L.F32)
                M32 = AMGPrecond(amg, A32, pinv, L.F32)
                A64r = A64f if mf else A64
                mixed_solve(bvec, A32, M32, A64r, 1e-8, maxiter=2000)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                xm, its, nref, rel = mixed_solve(bvec, A32, M32, A64r, 1e-8, maxiter=2000)
                torch.cuda.synchronize()
                dtm = time.perf_counter() - t0
                err = float(torch.linalg.norm(xm - xa) / torch.linalg.norm(xa))
                solves[name] = dict(inner_iters=its, refinements=nref, relres=rel, seconds=dtm,
                                    solves_per_s=1 / dtm, rel_err_vs_fp64_amg=err)
                log(f"[bench] {name}: {its} inner it, {nref} refinements, relres {rel:.2e}, "
                    f"{dtm * 1e3:.1f} ms, rel err vs FP64 {err:.2e}")
            # FP64 PCG preconditioned by the FP32 V-cycle (FP64 residual, dots, norms, coarse
            # solve); error against an FP64 AMG-PCG solve at tol 1e-12 (untimed)
            xt, _, _, _ = pcg_device(A64f, bvec, M64f, tol=1e-12, maxiter=4000)
            for name, A in (("amg_fp64_cycle32", A64), ("amg_fp64_cycle32_mf", A64f)):
                Mc = AMGPrecond(amg, A, pinv, L.F64, cycle=L.F32)
                pcg_device(A, bvec, Mc, tol=1e-8, maxiter=2000)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                xc, k, rr, st = pcg_device(A, bvec, Mc, tol=1e-8, maxiter=2000)
                torch.cuda.synchronize()
                dtc = time.perf_counter() - t0
                err = float(torch.linalg.norm(xc - xt) / torch.linalg.norm(xt))
                err64 = float(torch.linalg.norm(xa - xt) / torch.linalg.norm(xt))
                solves[name] = dict(iters=k, relres=rr, seconds=dtc, solves_per_s=1 / dtc,
                                    rel_err_vs_fp64_tol1e12=err, fp64_tol1e8_rel_err=err64)
                log(f"[bench] {name}: {k} it, relres {rr:.2e}, {dtc * 1e3:.1f} ms, rel err {err:.1e} "
                    f"(FP64 at the same tol: {err64:.1e})")
        extra["solves"] = solves

    if not args.quick and not args.no_multi:
        # 64 load cases on the same mesh: batched operator and block PCG (Jacobi, tol 1e-8)
        from exfem_amd.optimize import multi_load_cases
        from exfem_amd.solvers import Step
        Rn = 64
        F = workloads.top_face_loads(p, Rn)
        U = torch.as_tensor(np.random.default_rng(2).standard_normal((Rn, p.ndof)), device=dev)
        mres = {}
        for name, K, dt in (("fp64_sym", Ks, L.F64), ("fp32_sym", Ks32, L.F32)):
            op.set_ke(0, K)
            UU = U if dt == L.F64 else U.float()
            op.matvec_multi(UU, dtype=dt)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                op.matvec_multi(UU, dtype=dt)
            e1.record()
            torch.cuda.synchronize()
            tm = e0.elapsed_time(e1) / 5
            ss = 8 if dt == L.F64 else 4
            shape = op.multi_shape(Rn, dtype=dt)
            # the element-by-element MFMA product (slot variant, 4 case tiles per wavefront)
            # timed beside the tuned choice
            op.set_multi_shape(Rn, 4, dtype=dt)
            op.matvec_multi(UU, dtype=dt)
            e0.record()
            for _ in range(5):
                op.matvec_multi(UU, dtype=dt)
            e1.record()
            torch.cuda.synchronize()
            tmm = e0.elapsed_time(e1) / 5
            op.set_multi_shape(Rn, None, dtype=dt)
            op.matvec_multi(UU, dtype=dt)
            mres[name] = dict(ms_per_apply=tm, gdofs_per_rhs_total=Rn * p.ndof / (tm * 1e-3) / 1e9,
                              gflops=2 * p.ne * 576 * Rn / (tm * 1e-3) / 1e9,
                              algorithmic_bytes=p.ne * (300 * ss + 32) + 2 * Rn * p.ndof * ss,
                              tuned=shape, ebe_mfma_ms_per_apply=tmm)
            log(f"[bench] multi-RHS (64) matvec {name}: {tm:.2f} ms ({mres[name]['gdofs_per_rhs_total']:.1f} GDOF/s, "
                f"variant {shape['index']} nmax {shape['nmax']}); EbE MFMA slot variant {tmm:.2f} ms")
        fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
        step = Step(bcs=[(fix, d, 0.0) for d in range(3)])
        mres["block_pcg"] = {}
        if not args.no_block:
            # one FP64 reference for every block result: the AMG block PCG at tol 1e-12 on the
            # hierarchy built once here (its set-up timed apart)
            from exfem_amd.linalg import AMG
            from exfem_amd.solvers import State, _Setup
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            opa = m.operator()
            for i, K in enumerate(m.element_stiffness(dtype=L.F64)):
                opa.set_ke(i, K)
            opa.set_bc(_Setup(m, step, State(m)).free_u8)
            amg = AMG.from_operator(opa, m.X)
            amg_s = time.perf_counter() - t0
            Uref, _ = multi_load_cases(m, step, F, cg_tol=1e-12, return_info=True, precond="amg",
                                       amg=amg)
            rel_err = lambda Ub: float(((Ub.double() - Uref).norm(dim=1) / Uref.norm(dim=1)).max())
            runs = (("fp64", "jacobi", L.F64), ("fp32", "jacobi", L.F32),
                    ("amg_fp64", "amg", L.F64), ("amg_fp64_cycle32", "amg32", L.F64))
            for name, pc, dt in runs:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                Ub, info = multi_load_cases(m, step, F, cg_tol=1e-8, dtype=dt, return_info=True,
                                            precond=pc, amg=None if pc == "jacobi" else amg)
                torch.cuda.synchronize()
                dtb = time.perf_counter() - t0
                err = rel_err(Ub)
                rec = dict(cases=Rn, seconds=dtb, cases_per_s=Rn / dtb,
                           max_iters=int(info["iters"].max()), min_iters=int(info["iters"].min()),
                           max_relres=float(info["relres"].max()),
                           max_rel_err_vs_fp64_tol1e12=err)
                if pc != "jacobi":
                    rec["amg_setup_s"] = amg_s
                mres["block_pcg"][name] = rec
                log(f"[bench] block PCG 64 cases {name}: {dtb:.2f}s, iters "
                    f"{info['iters'].min()}-{info['iters'].max()}, max rel err vs FP64 {err:.1e}"
                    + (f" (AMG setup {amg_s:.2f}s apart)" if pc != "jacobi" else ""))
        extra["multi_rhs"] = mres

    if not args.quick and not args.no_serve:
        # the serving configuration: 512 requests of 1e4-2e6 dof (Poisson phase on the first
        # --serve-poisson-n of them)
        from exfem_amd import serving
        extra["serving"] = serving.benchmark(n=args.serve_n, dof_range=(1e4, args.serve_dof_max),
                                             log=log, progress=128, poisson_n=args.serve_poisson_n)
    if not args.quick and not args.no_tet10:
        # the precision configuration at its full size: the 42 x 42 x 168 C3D10 Delaunay box
        # (8.1M dof), mesh build and AMG set-up timed apart from the solves
        extra["tet10"] = tet10_workload(args, log, lite=True)

    cpu = None
    if not args.no_cpu:
        # CPU baseline: the C restatement of the reference's gather matvec (oracle/) on the
        # host's cores, FP64 full element matrices, a bounded sample of about cpu_seconds
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import cpu as ocpu
        op.set_ke(0, Ks)
        y = op.matvec(x)
        Kh = Kf.numpy()
        # the box's CPU share for one GPU is 16 cores (OMP_NUM_THREADS=16 there); the affinity
        # mask may show more cores than that share, so both counts are reported
        aff = len(os.sched_getaffinity(0))
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), aff, 16))
        import tempfile
        native_lib = ocpu.build(os.path.join(tempfile.mkdtemp(prefix="exfem_cpu_"), "liboracle_native.so"),
                                native=True)
        ocpu.load(native_lib)
        g = ocpu.GatherMatvec(p.conn, Kh, p.nn, threads=threads)
        xh = x.cpu().numpy()
        yo = g(xh)                                                    # warm-up
        err = float(np.linalg.norm(y.cpu().numpy() - yo) / np.linalg.norm(yo))
        reps = 0
        t0 = time.perf_counter()
        while reps < 10000 and (reps < 3 or time.perf_counter() - t0 < args.cpu_seconds):
            g(xh)
            reps += 1
        tc = (time.perf_counter() - t0) / reps
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        # solves/s of the CPU restatement of Jacobi-PCG (oracle_jacobi_pcg): config 0 repeated
        # for about 2 s, config 1 one full solve at tol 1e-8 (the GPU's element matrices as data)
        from oracle import gfem_oracle as O
        p0 = workloads.small_cantilever()
        K0 = O.element_stiffness(O.HEX8, p0.X[p0.conn], p0.E, p0.nu)
        g0 = ocpu.GatherMatvec(p0.conn, K0, p0.nn, threads=1)
        n0, t0 = 0, time.perf_counter()
        while n0 < 3 or time.perf_counter() - t0 < 2.0:
            _, k0, _ = g0.jacobi_pcg(p0.free * p0.f, p0.free, tol=1e-10)
            n0 += 1
        t_c0 = (time.perf_counter() - t0) / n0
        t0 = time.perf_counter()
        _, k1, r1 = g.jacobi_pcg(p.free * p.f, p.free, tol=1e-8)
        t_c1 = time.perf_counter() - t0
        cpu = dict(value=p.ndof / tc / 1e9, unit="GDOF/s", cores=threads, kind="port",
                   host_affinity_cores=aff, cpu_model=model or None,
                   sample=f"C restatement of gather_matvec (gfem/model.py:241-267; oracle/"
                          f"gather_matvec.c, gcc -O3 -march=native -fopenmp on this host), FP64 "
                          f"full Ke, same {p.ndof}-dof mesh and x, {reps} calls ({reps * tc:.1f} s) "
                          f"after 1 warm-up, {threads} OpenMP threads (the 16-core CPU share of one "
                          f"GPU; {aff} cores in the affinity mask) on {model or 'the host CPU'}",
                   seconds_per_matvec=tc, gpu_sym_vs_oracle_relerr=err,
                   jacobi_pcg=dict(
                       c0=dict(ndof=p0.ndof, tol=1e-10, iters=k0, seconds=t_c0, solves_per_s=1 / t_c0,
                               threads=1, sample=f"{n0} solves"),
                       c1=dict(ndof=p.ndof, tol=1e-8, iters=k1, relres=r1, seconds=t_c1,
                               solves_per_s=1 / t_c1, threads=threads, sample="1 solve")))
        log(f"[bench] CPU oracle ({threads} threads) {tc * 1e3:.1f} ms/matvec; GPU vs oracle rel err {err:.2e}; "
            f"Jacobi-PCG c0 {t_c0 * 1e3:.1f} ms ({k0} it, 1 thread), c1 {t_c1:.1f} s ({k1} it)")
        del Kh, g

    out = {
        "metric": "EbE matvec GDOF/s (% HBM peak); PCG solves/s at 1e-8 err; p50/p99 latency",
        "value": head["gdofs"],
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["us_per_matvec"] / 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, x ~ N(0,1) rng(1))",
        "config": {"workload": f"FP64 EbE matvec y = K x, C3D8 {nx}x{ny}x{nz} jittered cantilever "
                               f"box (ne {p.ne}, ndof {p.ndof}), upper-triangle FP64 element "
                               f"matrices", "ndof": p.ndof, "ne": p.ne},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
