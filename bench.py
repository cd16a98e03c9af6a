#!/usr/bin/env python3
"""Benchmark: SIMPLE iterations/s and cell-updates/s on the laminar
backward-facing step, 2000 x 500 quads (1M cells), on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision both|fp64|mixed] [--no-cpu]

A step is one SIMPLE loop iteration (Solvers_1_SIMPLE.jl:125-230): two momentum
assemblies + BiCGSTAB/DILU solves, the pressure equation (CG/DILU), gradients,
interpolation and the flux update, with every field resident in HBM.  Both
forms run by default: plain FP64 (the headline `value`) and mixed precision
(FP32 Krylov inner solves in FP64 defect correction).  Prints one JSON line.
The operators (DILU apply, the step's dominant kernel, and SpMV) are timed
live with HIP events on the library's stream; the CPU baseline is the oracle
(single-threaded FP64 restatement of the reference) on a bounded number of
iterations of the same case.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xcalibre_tw.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "SIMPLE iters/s & cell-updates/s at 1e-6 resid: mixed vs FP64; SpMV HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(nx, ny, iterations, precision):
    import fvmhip as F
    from tests.cases import bfs_model, config

    mesh = F.backward_facing_step(nx=nx, ny=ny)
    model = bfs_model(mesh)
    cfg = config(iterations, schemes=dict(divergence=F.Upwind), precision=precision)
    return mesh, model, cfg


def run_gpu(args, precision):
    import fvmhip as F
    from fvmhip import _lib as L
    from fvmhip import roofline as RL

    name = "fp64" if precision == L.FP64 else "mixed"
    mesh, model, cfg = make_case(args.nx, args.ny, args.warmup + args.steps + 1, precision)
    t0 = time.time()
    sim = F.Simulation(model, cfg, None, L.ALG_SIMPLE, conv_break=1e-6)
    log(f"[bench] {name}: set-up {time.time() - t0:.1f}s, {mesh.ncells} cells")
    done = sim.run(args.warmup)
    L.check(L.lib().fvh_synchronize())
    t1 = time.perf_counter()
    done = sim.run(args.steps)
    L.check(L.lib().fvh_synchronize())
    dt = time.perf_counter() - t1
    if done != args.steps:
        log(f"[bench] loop stopped after {done} timed iterations (converged)")
    R = sim.residuals()
    it0 = args.warmup
    res = dict(iters=done, seconds=dt, iters_per_s=done / dt, cell_updates_per_s=mesh.ncells * done / dt,
               ms_per_step=1e3 * dt / max(done, 1),
               residuals_last=[float(R[k][it0 + done - 1]) for k in range(3)],
               krylov=dict(Ux=sim.stats("Ux"), Uy=sim.stats("Uy"), p=sim.stats("p")))
    # operators on the assembled momentum matrix, timed with HIP events on the library stream
    eq = C.c_void_p()
    L.check(L.lib().fvh_sim_equation(sim.h, b"ux", C.byref(eq)))
    n = mesh.ncells
    rp, ci = mesh.device().pattern()
    nnz = int(ci.size)
    x = L.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n))
    y = L.DeviceArray(n)
    ms = C.c_double()

    def timed(op, reps, xv=x.vp, yv=y.vp):
        L.check(L.lib().fvh_time_op(eq, op, reps, xv, yv, C.byref(ms)))
        return ms.value

    ops = {}
    t = timed(0, 200)
    ops["spmv_fp64"] = dict(ms=t, bytes=RL.spmv(n, nnz), GBps=RL.gbps(RL.spmv(n, nnz), t))
    banded = mesh.device().sweep_info()[1] > 0
    if banded:  # the SpMV the solves use on a structured grid
        t = timed(6, 200)
        b = RL.spmv_band(n)
        ops["spmv_band_fp64"] = dict(ms=t, bytes=b, GBps=RL.gbps(b, t))
    t = timed(1, 20)
    ops["dilu_apply_fp64"] = dict(ms=t, bytes=RL.dilu_apply(n, nnz), GBps=RL.gbps(RL.dilu_apply(n, nnz), t))
    t = timed(2, 10, None, None)
    ops["dilu_update_fp64"] = dict(ms=t, bytes=RL.dilu_update(n, nnz), GBps=RL.gbps(RL.dilu_update(n, nnz), t))
    if precision != L.FP64:
        b = RL.spmv(n, nnz, 4, 4, 4)
        t = timed(3, 200)
        ops["spmv_fp32"] = dict(ms=t, bytes=b, GBps=RL.gbps(b, t))
        if banded:
            b = RL.spmv_band(n, 4, 4, 4)
            t = timed(7, 200)
            ops["spmv_band_fp32"] = dict(ms=t, bytes=b, GBps=RL.gbps(b, t))
        b = RL.dilu_apply(n, nnz, 4, 4, 4, 4, 4)
        t = timed(4, 20)
        ops["dilu_apply_fp32"] = dict(ms=t, bytes=b, GBps=RL.gbps(b, t))
    # loop phases on the simulation's own fields (HIP events, library stream)
    nif = mesh.nfaces - mesh.boundary_faces.size
    phases = {}
    for op, key, reps, nbytes in (
            (0, "assembly_ux", 50, RL.assembly(nif, nnz, n, k_f=4, k_c=3)),
            (1, "assembly_p", 50, RL.assembly(nif, nnz, n, k_f=3, k_c=2)),
            (2, "green_gauss_p", 100, RL.green_gauss(n, nif))):
        L.check(L.lib().fvh_sim_time_op(sim.h, op, reps, C.byref(ms)))
        phases[key] = dict(ms=ms.value, bytes=nbytes, GBps=RL.gbps(nbytes, ms.value),
                           frac=RL.gbps(nbytes, ms.value) / RL.HBM_PEAK_GBS)
    res["ops"] = ops
    res["phases"] = phases
    res["n"] = n
    res["nnz"] = nnz
    del sim
    return res


def run_spmv_large(args):
    """SpMV on a working set larger than the 256 MB Infinity Cache: a
    4000 x 1000 channel (4M cells, nnz 20M; CSR 336 MB with x and y), with
    synthetic diagonally dominant values, timed over back-to-back launches."""
    import fvmhip as F
    from fvmhip import _lib as L
    from fvmhip import roofline as RL

    nx, ny = args.spmv_large
    t0 = time.time()
    mesh = F.channel(nx, ny)
    dm = mesh.device()
    log(f"[bench] spmv-large: {mesh.ncells} cells, mesh {time.time() - t0:.1f}s")
    h = C.c_void_p()
    L.check(L.lib().fvh_eqn_create(dm.h, C.byref(h)))
    nz, b = C.c_void_p(), C.c_void_p()
    L.check(L.lib().fvh_eqn_arrays(h, C.byref(nz), C.byref(b)))
    rp, ci = dm.pattern()
    n, nnz = mesh.ncells, int(ci.size)
    rows = np.repeat(np.arange(n), np.diff(rp))
    vals = np.where(rows == ci, 4.5, -0.5 - 0.1 * np.random.default_rng(1).random(ci.size))
    L.check(L.lib().fvh_memcpy(nz, vals.ctypes.data_as(C.c_void_p), vals.nbytes))
    del rows, vals
    x = L.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n))
    y = L.DeviceArray(n)
    ms = C.c_double()
    out = dict(cells=n, nnz=nnz, workload=f"{nx}x{ny} channel")
    for op, key, nbytes in ((0, "csr_fp64", RL.spmv(n, nnz)), (6, "band_fp64", RL.spmv_band(n))):
        L.check(L.lib().fvh_time_op(h, op, 50, x.vp, y.vp, C.byref(ms)))
        out[key] = dict(ms=ms.value, bytes=nbytes, GBps=RL.gbps(nbytes, ms.value),
                        frac=RL.gbps(nbytes, ms.value) / RL.HBM_PEAK_GBS)
        log(f"[bench] spmv-large {key}: {ms.value * 1e3:.1f} us, {RL.gbps(nbytes, ms.value):.0f} GB/s "
            f"({nbytes / 1e6:.0f} MB)")
    L.lib().fvh_eqn_destroy(h)
    return out


def host_cpu():
    """Model name of the host CPU and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, avail


def run_to_break(args):
    """C2 run to the 1e-6 convergence break (Solvers_1_SIMPLE.jl:200-215) in
    both forms, capped at --max-iters: iterations to the break, time to
    solution, and the mixed form's per-field relative L2 error against the
    FP64 form."""
    import fvmhip as F
    from fvmhip import _lib as L
    from tests.cases import rel_l2

    out, fields = {}, {}
    for prec, name in ((L.FP64, "fp64"), (L.MIXED, "mixed")):
        mesh, model, cfg = make_case(args.nx, args.ny, args.max_iters, prec)
        sim = F.Simulation(model, cfg, None, L.ALG_SIMPLE, conv_break=1e-6)
        L.check(L.lib().fvh_synchronize())
        t1 = time.perf_counter()
        done, chunk = 0, 200
        while done < args.max_iters:
            k = sim.run(min(chunk, args.max_iters - done))
            done += k
            R = sim.residuals()
            last = [float(R[q][done - 1]) for q in range(3)]
            log(f"[bench] to-break {name}: {done} iterations, {time.perf_counter() - t1:.1f}s, residuals {last}")
            if k < chunk or max(last) <= 1e-6:
                break
        L.check(L.lib().fvh_synchronize())
        dt = time.perf_counter() - t1
        R = sim.residuals()
        last = [float(R[q][done - 1]) for q in range(3)]
        out[name] = dict(iterations=done, seconds=dt, converged=bool(max(last) <= 1e-6), residuals_last=last,
                         iters_per_s=done / dt, cell_updates_per_s=mesh.ncells * done / dt)
        fields[name] = {f: sim.get(f) for f in ("Ux", "Uy", "p")}
        del sim
    out["time_to_solution_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["rel_l2_mixed_vs_fp64"] = {f: rel_l2(fields["mixed"][f], fields["fp64"][f]) for f in ("Ux", "Uy", "p")}
    out["cells"] = args.nx * args.ny
    return out


def run_batch(args):
    """C5: 64 independent cavity/channel cases, sizes 32^2..512^2 (seeded
    log-uniform), Re 100..1000, advanced together (one launch per phase);
    against the same cases run one after another as single FP64 solves."""
    import fvmhip as F
    from fvmhip import _lib as L
    from tests.cases import batch_cases, config

    it = args.batch_iters
    out = {}
    cases = batch_cases(args.batch_cases)
    ncells = sum(m.mesh.ncells for m, _ in cases)
    for prec, name in ((L.FP64, "fp64"), (L.MIXED, "mixed")):
        cases = batch_cases(args.batch_cases)
        cfg = config(it + 1, precision=prec)
        t0 = time.time()
        b = F.BatchSimulation([m for m, _ in cases], cfg, [p for _, p in cases])
        setup = time.time() - t0
        b.run(1)  # warm-up iteration
        L.check(L.lib().fvh_synchronize())
        t1 = time.perf_counter()
        done = b.run(it)
        L.check(L.lib().fvh_synchronize())
        dt = time.perf_counter() - t1
        out[name] = dict(iters=done, seconds=dt, setup_s=setup, cell_updates_per_s=ncells * done / dt,
                         ms_per_iteration=1e3 * dt / done)
        log(f"[bench] batch {name}: {1e3 * dt / done:.1f} ms/iteration over {len(cases)} cases ({ncells} cells)")
        del b
    # the same cases, one after another, single-case FP64
    cases = batch_cases(args.batch_cases)
    cfg = config(it + 1)
    total = 0.0
    for m, p in cases:
        s = F.Simulation(m, cfg, p)
        s.run(1)
        L.check(L.lib().fvh_synchronize())
        t1 = time.perf_counter()
        s.run(it)
        L.check(L.lib().fvh_synchronize())
        total += time.perf_counter() - t1
        del s
    out["sequential_fp64"] = dict(iters=it, seconds=total, cell_updates_per_s=ncells * it / total,
                                  ms_per_iteration=1e3 * total / it)
    log(f"[bench] sequential fp64: {1e3 * total / it:.1f} ms per iteration of all cases")
    out["cases"] = len(cases)
    out["cells"] = ncells
    out["speedup_vs_sequential_fp64"] = out["mixed"]["cell_updates_per_s"] / out["sequential_fp64"]["cell_updates_per_s"]
    return out


def run_c3(args):
    """C3: steady convection-diffusion on a Delaunay triangulation of seeded
    uniform points (1M by default, ~2M triangles), u = (sin pi x cos pi y,
    -cos pi x sin pi y), gamma = 1e-3, phi = 1 on x = 0 and 0 elsewhere,
    upwind convection; one BiCGSTAB + DILU solve to rtol 1e-8 (at most 500
    iterations: BiCGSTAB + DILU needs thousands at 2M cells, so the figure of
    merit is the time per Krylov iteration), FP64 and FP32-in-FP64."""
    import fvmhip as F
    from fvmhip import _lib as L
    from tests.cases import rel_l2

    t0 = time.time()
    mesh = F.delaunay_unit_square(args.c3_points, seed=0)
    log(f"[bench] c3: mesh {mesh.ncells} cells in {time.time() - t0:.1f}s")
    bcs = [F.Dirichlet("left", 1.0), F.Dirichlet("right", 0.0), F.Dirichlet("bottom", 0.0), F.Dirichlet("top", 0.0)]
    out = dict(cells=mesh.ncells, levels=mesh.device().sweep_info()[0])
    sol = {}
    for prec, name in ((L.FP64, "fp64"), (L.MIXED, "mixed")):
        eq = F.ScalarEquation(mesh, F.taylor_green_flux(mesh), np.full(mesh.nfaces, 1e-3), bcs, divergence=F.Upwind)
        L.check(L.lib().fvh_synchronize())
        t1 = time.perf_counter()
        eq.discretise()
        L.check(L.lib().fvh_synchronize())
        t_asm = time.perf_counter() - t1
        t1 = time.perf_counter()
        st = eq.solve(rtol=1e-8, itmax=500, precision=prec)
        L.check(L.lib().fvh_synchronize())
        dt = time.perf_counter() - t1
        sol[name] = eq.values()
        out[name] = dict(assembly_ms=1e3 * t_asm, solve_ms=1e3 * dt, niter=st["niter"], status=st["status"],
                         refinements=st["refinements"], rnorm=st["rnorm"], rnorm0=st["rnorm0"],
                         ms_per_krylov_iter=1e3 * dt / max(st["niter"], 1))
        log(f"[bench] c3 {name}: solve {1e3 * dt:.0f} ms, {st['niter']} iterations ({st['status']})")
        del eq
    out["mixed_vs_fp64"] = out["fp64"]["solve_ms"] / out["mixed"]["solve_ms"]
    out["rel_l2_mixed_vs_fp64"] = rel_l2(sol["mixed"], sol["fp64"])
    return out


def run_c4(args):
    """C4: k-omega flat plate, 1000 x 500 quads graded 1.05 from the wall,
    U = 10 m/s, nu = 1.5e-5, k / omega from 1 % intensity and viscosity ratio
    10; five solves per SIMPLE iteration; FP64 and mixed."""
    import fvmhip as F
    from fvmhip import _lib as L
    from tests.cases import config, flatplate_model

    out = {}
    for prec, name in ((L.FP64, "fp64"), (L.MIXED, "mixed")):
        mesh = F.flat_plate(nx=1000, ny=500, ratio=1.05)
        model = flatplate_model(mesh)
        cfg = config(args.warmup + args.steps + 1, schemes=dict(divergence=F.Upwind), turbulent=True,
                     p_solver=F.CgSolver, precision=prec)
        sim = F.Simulation(model, cfg, None, L.ALG_SIMPLE, conv_break=1e-6)
        sim.run(args.warmup)
        L.check(L.lib().fvh_synchronize())
        t1 = time.perf_counter()
        done = sim.run(args.steps)
        L.check(L.lib().fvh_synchronize())
        dt = time.perf_counter() - t1
        out[name] = dict(cells=mesh.ncells, iters=done, iters_per_s=done / dt, cell_updates_per_s=mesh.ncells * done / dt,
                         ms_per_step=1e3 * dt / max(done, 1),
                         krylov={f: sim.stats(f) for f in ("Ux", "p", "k", "omega")})
        log(f"[bench] c4 {name}: {done / dt:.2f} it/s")
        del sim
    out["mixed_vs_fp64"] = out["mixed"]["iters_per_s"] / out["fp64"]["iters_per_s"]
    return out


def run_cpu(args):
    """The oracle (1 core) on the same case: a bounded sample of iterations."""
    import fvmhip as F
    from oracle import ref

    mesh, model, cfg = make_case(args.nx, args.ny, args.cpu_iters + 1, 0)
    case = F.CaseDesc(model, cfg, None, 0, 1e-6)
    t0 = time.time()
    rs = ref.RefSimulation(model, case)
    setup = time.time() - t0
    t1 = time.perf_counter()
    n = rs.run(args.cpu_iters)
    dt = time.perf_counter() - t1
    return dict(iters=n, seconds=dt, setup_s=setup, cell_updates_per_s=mesh.ncells * n / dt, iters_per_s=n / dt)


def profile_ops(args):
    """One launch of each operator on a C2-pattern matrix (diagonally dominant
    synthetic values): the plain copy (reference for traffic counters), SpMV
    (CSR, and banded on a structured grid) and the DILU apply in FP64 and
    FP32.  Few kernels, for counter passes."""
    import fvmhip as F
    from fvmhip import _lib as L

    mesh = F.backward_facing_step(nx=args.nx, ny=args.ny)
    dm = mesh.device()
    h = C.c_void_p()
    L.check(L.lib().fvh_eqn_create(dm.h, C.byref(h)))
    nz, b = C.c_void_p(), C.c_void_p()
    L.check(L.lib().fvh_eqn_arrays(h, C.byref(nz), C.byref(b)))
    rp, ci = dm.pattern()
    rows = np.repeat(np.arange(mesh.ncells), np.diff(rp))
    vals = np.where(rows == ci, 4.5, -0.5 - 0.1 * np.random.default_rng(1).random(ci.size))
    L.check(L.lib().fvh_memcpy(nz, vals.ctypes.data_as(C.c_void_p), vals.nbytes))
    L.check(L.lib().fvh_precond_set(h, L.PRECOND_DILU, None))
    L.check(L.lib().fvh_precond_update(h))
    n = mesh.ncells
    x = L.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n))
    y = L.DeviceArray(n)
    ms = C.c_double()
    ops = (5, 0, 1, 3, 4) + ((6, 7) if dm.sweep_info()[1] > 0 else ())
    for op in ops:
        L.check(L.lib().fvh_time_op(h, op, 1, x.vp, y.vp, C.byref(ms)))
        log(f"[profile-ops] op {op}: {ms.value * 1e3:.1f} us")
    L.check(L.lib().fvh_synchronize())
    L.lib().fvh_eqn_destroy(h)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--nx", type=int, default=2000)
    ap.add_argument("--ny", type=int, default=500)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-iters", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4"],
                    help="c2: the headline BFS (default); c3: Delaunay transport solve; c4: k-omega flat plate")
    ap.add_argument("--c3-points", type=int, default=1_000_000)
    ap.add_argument("--spmv-large", type=int, nargs=2, default=[4000, 1000], metavar=("NX", "NY"),
                    help="grid of the SpMV leg whose working set exceeds the Infinity Cache")
    ap.add_argument("--no-spmv-large", action="store_true")
    ap.add_argument("--to-break", action="store_true",
                    help="run C2 to the 1e-6 break in both forms (minutes) and print that record instead")
    ap.add_argument("--max-iters", type=int, default=6000)
    ap.add_argument("--profile-ops", action="store_true",
                    help="run each operator once on the C2 case (copy, SpMV, DILU apply; FP64 and FP32), for counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU decomposition is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import fvmhip as F
    from fvmhip import _lib as L
    from fvmhip import roofline as RL

    if F.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.profile_ops:
        profile_ops(args)
        return
    if args.workload != "c2":
        res = run_c3(args) if args.workload == "c3" else run_c4(args)
        print(json.dumps({"workload": args.workload, **res}))
        return
    if args.to_break:
        print(json.dumps({"workload": "c2-to-break", **run_to_break(args)}))
        return
    forms = {"fp64": [L.FP64], "mixed": [L.MIXED], "both": [L.FP64, L.MIXED]}[args.precision]
    out = {}
    for p in forms:
        k = "fp64" if p == L.FP64 else "mixed"
        out[k] = run_gpu(args, p)
        log(f"[bench] {k}: {out[k]['iters_per_s']:.3f} it/s, {out[k]['ms_per_step']:.1f} ms/iteration")
    # headline: the FP64 form (the reference's arithmetic) when it ran
    hk = "fp64" if "fp64" in out else "mixed"
    head = out[hk]
    cpu = None
    if not args.no_cpu:
        c = run_cpu(args)
        model, avail = host_cpu()
        cpu = dict(value=c["cell_updates_per_s"], unit="cell-updates/s", cores=1, kind="port",
                   host_cpu=model, host_cores_available=avail,
                   sample=f"{c['iters']} SIMPLE iterations of the same {args.nx}x{args.ny} case, oracle FP64 "
                          f"restatement, {c['seconds']:.1f}s (set-up {c['setup_s']:.1f}s excluded), "
                          f"1 of {avail} cores of a {model}")
        log(f"[bench] cpu: {c['iters_per_s']:.4f} it/s")
    # dominant kernel of the step: the DILU apply (banded sweep), called by every Krylov iteration
    dk = "dilu_apply_fp64" if hk == "fp64" else "dilu_apply_fp32"
    d = head["ops"][dk]
    # HBM bytes per launch measured by rocprofv3 counters on this workload
    # (FETCH_SIZE / WRITE_SIZE passes of `bench.py --profile-ops`, scaled by a
    # plain copy of known size); null when the summary is absent
    traffic = None
    tf = os.path.join(ROOT, "profiles", "c2_operators_hbm_traffic.json")
    if os.path.exists(tf) and (args.nx, args.ny) == (2000, 500):
        with open(tf) as f:
            t = json.load(f)
        traffic = t["dilu_apply_fp64_bytes" if hk == "fp64" else "dilu_apply_fp32_bytes"]
    sp = head["ops"]["spmv_fp64"]
    line = {
        "metric": METRIC,
        "value": head["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if hk == "fp64" else "f32-in-f64",
        "data": "synthetic (generated BFS mesh, zero-pressure / uniform-velocity initial fields)",
        "config": {"workload": f"laminar backward-facing step {args.nx}x{args.ny} quads ({head['n']} cells), "
                               f"SIMPLE, Upwind convection, BiCGSTAB+DILU (U), CG+DILU (p), Re_h=500",
                   "cells": head["n"], "nnz": head["nnz"], "form": hk},
        "simple_iters_per_s": head["iters_per_s"],
        "roofline": {"bound": "hbm", "achieved": d["GBps"], "peak": RL.HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": d["GBps"] / RL.HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("rocprofv3 FETCH_SIZE / WRITE_SIZE passes of bench.py --profile-ops, "
                                        "committed in profiles/c2_operators_hbm_traffic.json (not measured in "
                                        "this run)") if traffic is not None else None,
                     "kernel": "DILU apply: k_skew_gather + k_skew_fwd + k_skew_bwd + k_skew_scatter "
                               "(pipelined anti-diagonal sweeps)",
                     "bytes_per_launch": d["bytes"], "ms_per_launch": d["ms"]},
        "spmv": {"kernel": "k_spmv (FP64 CSR)", "GBps": sp["GBps"], "frac": sp["GBps"] / RL.HBM_PEAK_GBS,
                 "bytes_per_launch": sp["bytes"], "ms_per_launch": sp["ms"]},
        "spmv_band": ({"kernel": "k_spmv_band (FP64, A by band position: the SpMV of the solves on this grid)",
                       "GBps": head["ops"]["spmv_band_fp64"]["GBps"],
                       "frac": head["ops"]["spmv_band_fp64"]["GBps"] / RL.HBM_PEAK_GBS,
                       "bytes_per_launch": head["ops"]["spmv_band_fp64"]["bytes"],
                       "ms_per_launch": head["ops"]["spmv_band_fp64"]["ms"]}
                      if "spmv_band_fp64" in head["ops"] else None),
        "phases": {k: {q: head["phases"][k][q] for q in ("GBps", "frac", "ms", "bytes")} for k in head["phases"]},
        "cpu_baseline": cpu,
        "forms": out,
    }
    if not args.no_spmv_large:
        line["spmv_large"] = run_spmv_large(args)
    if "fp64" in out and "mixed" in out:
        line["mixed_vs_fp64"] = out["mixed"]["iters_per_s"] / out["fp64"]["iters_per_s"]
    if not args.no_batch:
        line["batch"] = run_batch(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
