"""Benchmark of the FP64 EbE stiffness matvec on one MI355X.

Workload: C3D8 200x40x40 cantilever box of 100x20x20 with interior nodes jittered by
+-10% of the element size (seed 0): 320,000 elements, 337,881 nodes, 1,013,643 dof. One step =
one operator application y = K x over the whole mesh (element pass streaming the packed FP64
element matrices + deterministic node-gather assembly), inputs resident in HBM.

Prints one JSON line: GDOF/s of the matvec as `value`, the element-pass roofline (algorithmic
bytes / HIP-event duration vs 8 TB/s), and the CPU baseline (the oracle's NumPy gather matvec,
single thread, on a bounded number of repetitions of the same workload).

    python bench.py [--steps K] [--warmup W] [--storage full|sym] [--pcg] [--no-cpu]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--storage", choices=["full", "sym"], default="full")
    ap.add_argument("--nx", type=int, nargs=3, default=[200, 40, 40])
    ap.add_argument("--pcg", action="store_true", help="also time one Jacobi-PCG solve (tol 1e-8)")
    ap.add_argument("--amg", action="store_true", help="also time one AMG-PCG solve (tol 1e-8)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=20)
    ap.add_argument("--calib", action="store_true",
                    help="also run one plain 256 MiB device copy (counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the operator would "
              "shard by element slabs with a halo exchange); run it with --gpus 1", flush=True)
        sys.exit(2)

    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    from exfem_amd.model import stream_ptr, vptr

    L.load()
    dev = torch.device("cuda")
    nx, ny, nz = args.nx
    t0 = time.time()
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=0)
    log(f"[bench] mesh {p.name}: ne={p.ne} nn={p.nn} ndof={p.ndof} ({time.time() - t0:.1f}s)")
    m = p.model()
    storage = L.KE_SYM if args.storage == "sym" else L.KE_FULL
    torch.cuda.synchronize()
    t0 = time.time()
    Ke = m.element_stiffness(storage=storage)[0]
    torch.cuda.synchronize()
    ke_ms = (time.time() - t0) * 1e3
    op = m.operator()
    op.set_ke(0, Ke)
    x = torch.as_tensor(np.random.default_rng(1).standard_normal(p.ndof), device=dev)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        op.matvec(x, y)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        op.matvec(x, y)
    ev1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = ev0.elapsed_time(ev1) / args.steps
    log(f"[bench] matvec {ms * 1e3:.1f} us/step (wall {t_wall / args.steps * 1e6:.1f} us)")

    if args.calib:
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # per-pass HIP events on the launch stream
    te, tn = C.c_float(), C.c_float()
    nrep = max(10, args.steps // 2)
    L.check(L.load().exfem_op_profile(op.h, vptr(x), vptr(y), L.F64, 0, 0, nrep, C.byref(te),
                                      C.byref(tn), stream_ptr()))
    s = 8
    nent = 576 if storage == L.KE_FULL else 300
    b_elem = p.ne * (nent * s + 8 * 4) + p.ndof * s                   # Ke + conn + x
    b_matvec = p.ne * (nent * s + 8 * 4) + 2 * p.ndof * s               # + y written
    achieved = b_elem / (te.value * 1e-3)
    roofline = dict(bound="hbm", kernel="ebe_elem_kernel<double,8,%s>" % ("sym" if storage else "full"),
                    achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                    frac=achieved / HBM_PEAK, traffic=None,
                    elem_pass_us=te.value * 1e3, node_pass_us=tn.value * 1e3,
                    algorithmic_bytes=b_elem)
    gdofs = p.ndof / (ms * 1e-3) / 1e9
    extra = dict(matvec_bytes=b_matvec, matvec_hbm_frac=b_matvec / (ms * 1e-3) / HBM_PEAK,
                 ke_compute_ms=ke_ms, wall_us_per_step=t_wall / args.steps * 1e6)

    if args.pcg or args.amg:
        from exfem_amd.linalg import BcOperator, Jacobi, pcg_device
        op.set_bc(p.free)
        free = torch.as_tensor(p.free.astype(np.float64), device=dev)
        d = op.diagonal()
        pinv = torch.where(free > 0, 1.0 / torch.where(d.abs() > 0, d, torch.ones_like(d)), torch.ones_like(d))
        b = torch.as_tensor(p.free * p.f, device=dev)
        if args.pcg:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs, k, rr, st = pcg_device(BcOperator(op), b, Jacobi(pinv), tol=1e-8, maxiter=20000)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            extra.update(jacobi_pcg=dict(iters=k, relres=rr, seconds=dt, solves_per_s=1 / dt,
                                         ms_per_iter=dt / max(k, 1) * 1e3))
            log(f"[bench] Jacobi-PCG {k} it, relres {rr:.2e}, {dt:.2f}s")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import gfem_oracle as O
        Kh = Ke.numpy() if storage == L.KE_FULL else Ke.numpy()
        dofs = O.dofs(p.conn)
        plan = O.gather_plan([p.conn], p.nn)
        xh = x.cpu().numpy()
        yo = O.gather_matvec(plan, [(dofs, Kh)], xh, p.ndof)        # warm-up
        err = float(np.linalg.norm(y.cpu().numpy() - yo) / np.linalg.norm(yo))
        reps = args.cpu_reps
        t0 = time.perf_counter()
        for _ in range(reps):
            O.gather_matvec(plan, [(dofs, Kh)], xh, p.ndof)
        tc = (time.perf_counter() - t0) / reps
        cpu = dict(value=p.ndof / tc / 1e9, unit="GDOF/s", cores=1, kind="port",
                   sample=f"oracle NumPy gather_matvec (gfem/model.py:264-267 restated), FP64, full "
                          f"{p.ndof}-dof mesh, mean of {reps} calls after 1 warm-up, 1 thread",
                   seconds_per_matvec=tc, gpu_vs_oracle_relerr=err)
        log(f"[bench] CPU oracle {tc * 1e3:.0f} ms/matvec; GPU vs oracle rel err {err:.2e}")
        del Kh

    out = {
        "metric": "EbE matvec GDOF/s (% HBM peak); PCG solves/s at 1e-8 err; p50/p99 latency",
        "value": gdofs,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, x ~ N(0,1) rng(1))",
        "config": {"workload": f"FP64 EbE matvec, C3D8 {nx}x{ny}x{nz} jittered cantilever box "
                               f"(ne {p.ne}, ndof {p.ndof}), Ke storage {args.storage}",
                   "ndof": p.ndof, "ne": p.ne},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
