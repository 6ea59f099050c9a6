#!/usr/bin/env python3
"""Benchmark: FP64 Poisson Ax GDOF/s and p-multigrid PCG time-to-residual on
one MI355X, on the deformed-cube workload (32^3 hexahedra at polynomial
order 7, i.e. 8 GLL points per direction, 16.8M local nodes, Dirichlet
walls, x' = x + 0.1 sin(pi x) sin(pi y) sin(pi z) on each coordinate,
seeded 16-mode Fourier source, p-MG orders 7->3->1, tolerance 1e-10
relative in the multiplicity-weighted norm).

A "step" is one application of the FP64 Poisson operator (ssx_ax) to the
whole mesh; `value` is local DOFs per second over the timed steps, inputs
resident in HBM.  The full FCG + p-MG solve is timed once after the steps
and reported in the same JSON line.  HIP events are recorded on the stream
the kernels run on (the library is bound to torch's current stream).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_PER_NODE_POISSON = 64      # u 8 + Au 8 + six factors G*w*detJ 48
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles",
                               "ax_poisson_n8_cube32_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per launch of the Poisson Ax kernel measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (calibrated), if profiled."""
    try:
        for line in open(PROFILE_TRAFFIC):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def smooth_source(x, y, z, seed=2024):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 9, size=(16, 3))
    a = rng.standard_normal(16)
    f = np.zeros_like(x)
    for m in range(16):
        f += a[m] * (np.sin(np.pi * k[m, 0] * x) * np.sin(np.pi * k[m, 1] * y)
                     * np.sin(np.pi * k[m, 2] * z))
    return f


def solve_case(S, torch, ctx, mesh, max_iter, precs=("fp64", "fp32", "bf16"),
               log_prefix=""):
    """FCG + p-MG Dirichlet Poisson solve of the seeded Fourier source to
    1e-10 relative (multiplicity-weighted norm), once per smoother
    precision; the FP64 solution is the reference for the others."""
    f = smooth_source(mesh.x.ravel(), mesh.y.ravel(), mesh.z.ravel())
    massv = torch.from_numpy(ctx.level_data(0, S.DATA_MASS)).cuda()
    b0 = torch.from_numpy(f).cuda() * massv
    ctx.dss(0, b0)
    ctx.mask(0, S.X_VELOCITY, b0)
    bnorm = np.sqrt(ctx.dot(0, b0, b0))
    eps = 1e-10 * bnorm
    ctx.set_diag_choice(0)       # Dirichlet Poisson: Poisson diagonal
    code = {"fp64": S.PREC_FP64, "fp32": S.PREC_FP32, "bf16": S.PREC_BF16}
    ref_sol, out = None, {}
    for name in precs:
        prec = code[name]
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        if prec != S.PREC_FP64:   # build (and tune) the reduced copies
            w = torch.zeros_like(b0)
            ctx.mg_precon(S.OP_POISSON, S.X_VELOCITY, eps, b0.clone(), w)
        b = b0.clone()
        sol = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, b, sol, eps,
                     max_iter, allow_notconv=True)
        torch.cuda.synchronize()
        tsolve = time.perf_counter() - t0
        rec = {"time_to_residual_s": tsolve, "iterations": st.iterations,
               "converged": bool(st.converged),
               "final_residual_rel": st.residual / bnorm,
               "tolerance_rel": 1e-10, "smoother_precision": name,
               "outer_precision": "fp64", "fine_ax": st.fine_ax,
               "coarse_cg_iters": st.coarse_iters,
               "smoother_cg_iters": st.smooth_iters, "vcycles": st.vcycles}
        if ref_sol is None:
            ref_sol = sol
        else:
            d = sol - ref_sol
            rec["error_vs_fp64_rel"] = float(np.sqrt(
                ctx.dot(0, d, d) / ctx.dot(0, ref_sol, ref_sol)))
            rec["speedup_vs_fp64"] = (out["fp64"]["time_to_residual_s"]
                                      / tsolve)
        out[name] = rec
        log(f"[bench] {log_prefix}solve {name} {tsolve:.3f} s, "
            f"{st.iterations} its")
    ctx.set_option(S.OPT_SMOOTHER_PRECISION, S.PREC_FP64)
    return out


def particle_case(S, torch, stream, npart=1_000_000, seed=7):
    """Inertial point particles (integrate_particle_equations), and RK2
    tracers (ssx_tracers_step) from the same positions, on the
    channel workload: 24x16x24 hexahedra at n = 10 on
    [0,4pi] x [-1,1] x [0,4pi/3], tanh wall stretching 2.0, periodic in x
    and z, walls in y.  Velocity: Poiseuille u = 1 - y^2 plus a seeded
    divergence-free perturbation of 10% amplitude, set three times (the
    3-level history).  npart particles at uniform random element and
    reference position, velocity 0, response time tau ~ U[0.05, 0.5],
    advanced over one step dt = 2e-3 (one element is about 0.2 long in x,
    so few particles change element)."""
    from ssenss_amd.mesh import box_mesh, PERIODIC, NOSLIP
    mesh = box_mesh(10, (24, 16, 24), lo=(0.0, -1.0, 0.0),
                    hi=(4 * np.pi, 1.0, 4 * np.pi / 3),
                    bc=(PERIODIC, PERIODIC, NOSLIP, NOSLIP, PERIODIC,
                        PERIODIC), deform=2, amp=2.0)
    ctx = S.Context(mesh)
    ctx.set_stream(stream)
    X, Y, Z = (a.ravel() for a in (mesh.x, mesh.y, mesh.z))
    kx, kz = 2 * np.pi / (4 * np.pi), 2 * np.pi / (4 * np.pi / 3)
    pert = 0.1 * (1 - Y ** 2)
    u = (1 - Y ** 2) + pert * np.cos(kx * X) * np.sin(kz * Z)
    v = 0.1 * (1 - Y ** 2) ** 2 * np.sin(kx * X) * np.sin(kz * Z)
    w = -pert * np.sin(kx * X) * np.cos(kz * Z)
    uv = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (u, v, w)]
    for _ in range(3):
        ctx.particles_set_velocity(*uv)
    rng = np.random.default_rng(seed)
    dt, T = 2e-3, 1.0
    ip = lambda a: torch.from_numpy(a.astype(np.int32)).cuda()
    fp = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    el = rng.integers(0, mesh.nel, npart)
    xyz = rng.uniform(-1, 1, (3, npart))
    zero = np.zeros(npart)
    base = {"el": el, "xyz": xyz, "tau": rng.uniform(0.05, 0.5, npart)}

    def fresh():
        return {"el": ip(base["el"]), "old_el": ip(base["el"]),
                "x": fp(base["xyz"][0]), "y": fp(base["xyz"][1]),
                "z": fp(base["xyz"][2]), "u": fp(zero), "v": fp(zero),
                "w": fp(zero), "t": fp(zero + (T - dt)),
                "tau": fp(base["tau"]), "reinit": ip(np.zeros(npart))}
    ctx.particles_step(fresh(), T, dt)              # warm-up
    times = []
    for _ in range(3):
        pr = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        passes = ctx.particles_step(pr, T, dt)      # synchronous
        times.append(time.perf_counter() - t0)
    moved = int((pr["el"].cpu().numpy() != base["el"]).sum())
    reinit = int(pr["reinit"].sum().item())
    # RK2 tracers (the channel workload's tracer mode) on the same field and
    # positions, one step of dt
    ctx.tracers_step(fresh(), T, dt)                # warm-up
    tt = []
    for _ in range(3):
        tr = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tpasses = ctx.tracers_step(tr, T, dt)      # synchronous
        tt.append(time.perf_counter() - t0)
    tmoved = int((tr["el"].cpu().numpy() != base["el"]).sum())
    treinit = int(tr["reinit"].sum().item())
    ctx.close()
    t = min(times)
    return {"particles": npart, "local_dofs": mesh.ntot, "seconds": t,
            "particles_per_s": npart / t, "max_passes": passes,
            "changed_element": moved, "reseeded": reinit,
            "mesh": "channel 24x16x24, n=10, tanh 2.0", "dt": dt,
            "tracers_rk2": {"tracers": npart, "seconds": min(tt),
                            "tracers_per_s": npart / min(tt),
                            "max_passes": tpasses,
                            "changed_element": tmoved,
                            "reseeded": treinit}}


def batch_case(S, torch, nprob=16, seed=11, mrange=(8, 16),
               points=(5, 6, 8, 10, 12)):
    """Uneven batch (the ensemble workload, reduced to nprob problems
    of at most 16^3 elements so the bench stays within minutes):
    independent Helmholtz problems on randomly deformed boxes, element
    counts m^3 with m in mrange, orders N in {4,5,7,9,11} (points per
    direction n = N + 1 from `points`), lambda ~ U[1, 100], Jacobi-PCG to
    1e-10 relative.  The batch (one launch per CG stage, graph-replayed)
    against the same problems solved one by one; both are timed on their
    second run."""
    from ssenss_amd.mesh import box_mesh
    rng = np.random.default_rng(seed)
    ctxs, rhs = [], []
    for _ in range(nprob):
        m = int(rng.integers(mrange[0], mrange[1] + 1))
        n = int(rng.choice(list(points)))
        lam = rng.uniform(1, 100)
        mesh = box_mesh(n, (m, m, m), deform=1, amp=rng.uniform(0, 0.15),
                        dt=(11.0 / 6.0) / lam, visc=1.0)
        c = S.Context(mesh)
        c.set_stream(torch.cuda.current_stream())
        f = smooth_source(mesh.x.ravel(), mesh.y.ravel(), mesh.z.ravel(),
                          seed=int(rng.integers(1 << 30)))
        massv = torch.from_numpy(c.level_data(0, S.DATA_MASS)).cuda()
        b = torch.from_numpy(f).cuda() * massv
        c.dss(0, b)
        ctxs.append(c)
        rhs.append(b)
    eps = [1e-10 * float(np.sqrt(c.dot(0, b, b))) for c, b in zip(ctxs, rhs)]
    dofs = sum(c.ntot(0) for c in ctxs)
    for rep in range(2):                 # the first run is a warm-up
        fs = [b.clone() for b in rhs]
        us = [torch.zeros_like(f) for f in fs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        single_its = []
        for c, f, u, e in zip(ctxs, fs, us, eps):
            st = c.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, e, 5000)
            single_its.append(st.iterations)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
    batch = S.Batch([(c, S.OP_HELMHOLTZ, S.X_VELOCITY) for c in ctxs])
    res = {}
    for graph in (False, True):
        fs = [b.clone() for b in rhs]
        us = [torch.zeros_like(f) for f in fs]
        batch.cg(fs, us, eps, 5000, use_graph=graph)        # warm-up
        fs = [b.clone() for b in rhs]
        us = [torch.zeros_like(f) for f in fs]
        t0 = time.perf_counter()
        st = batch.cg(fs, us, eps, 5000, use_graph=graph)
        res[graph] = (time.perf_counter() - t0, [s.iterations for s in st])
    batch.close()
    for c in ctxs:
        c.close()
    return {"problems": nprob, "local_dofs": dofs,
            "elements_per_dir": list(mrange),
            "orders": sorted({c.mesh.n - 1 for c in ctxs}),
            "sequential_s": t_seq, "batched_s": res[False][0],
            "batched_graph_s": res[True][0],
            "speedup_batched_graph": t_seq / res[True][0],
            "iterations_single": single_its,
            "iterations_batched": res[True][1],
            "max_iteration_difference": max(
                abs(a - b) for a, b in zip(single_its, res[True][1]))}


def cpu_baseline(mesh, seconds):
    """The oracle (C restatement of the reference CPU path) on a bounded
    sample: Poisson Ax on a 12^3-element piece of the same deformed cube
    at the same order, repeated for about `seconds` of wall time."""
    from oracle import pyoracle as P
    from ssenss_amd.mesh import box_mesh
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    P.set_threads(threads)
    sm = box_mesh(mesh.n, (12, 12, 12), deform=1, amp=0.1)
    o = P.Oracle(sm)
    u = np.random.default_rng(1).standard_normal(sm.ntot)
    o.ax(0, P.OP_POISSON, u)
    reps, t0 = 0, time.perf_counter()
    while True:
        o.ax(0, P.OP_POISSON, u)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    o.close()
    return {"value": reps * sm.ntot / el / 1e9, "unit": "GDOF/s",
            "cores": P.get_threads(), "kind": "port",
            "sample": f"FP64 Poisson Ax, 12^3 elements at n={mesh.n} "
                      f"({sm.ntot} local nodes) x {reps} applications, "
                      f"{el:.1f} s, OpenMP over elements"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--elems", type=int, default=32)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-particles", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the N = 5 / 7 / 11 mixed-precision sweep")
    ap.add_argument("--profile", action="store_true",
                    help="kernels only, for rocprofv3 counter runs: the "
                    "Poisson Ax on the workload, then its DSS, two outer "
                    "FCG iterations with FP64 and FP32 smoothers, then a 1 GiB "
                    "device copy (known bytes, to calibrate FETCH_SIZE/"
                    "WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this "
              "engine runs all elements on a single MI355X and has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import ssenss_amd as S
    from ssenss_amd.mesh import box_mesh

    m = args.elems
    log(f"[bench] building {m}^3 mesh n={args.n}")
    mesh = box_mesh(args.n, (m, m, m), deform=1, amp=0.1)
    t0 = time.perf_counter()
    ctx = S.Context(mesh)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream)
    log(f"[bench] context set-up {time.perf_counter() - t0:.1f} s")
    nt = mesh.ntot
    rng = np.random.default_rng(0)
    u = torch.from_numpy(rng.standard_normal(nt)).cuda()
    au = torch.zeros_like(u)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    if args.profile:
        for _ in range(args.steps):
            ctx.ax(0, S.OP_POISSON, u, au)
        for _ in range(args.steps):
            ctx.dss(0, au)
        # two outer FCG iterations per smoother precision: the fused CG
        # update with the gather DSS, the residual kernels and the
        # coarse-level CG
        solve_case(S, torch, ctx, mesh, 2, precs=("fp64", "fp32"),
                   log_prefix="profile ")
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)                                   # 1 GiB each way
        torch.cuda.synchronize()
        log("[bench] profile mode done")
        ctx.close()
        return
    t_ax = timed(lambda: ctx.ax(0, S.OP_POISSON, u, au), args.steps,
                 args.warmup)
    log(f"[bench] Ax {t_ax * 1e6:.1f} us")
    t_helm = timed(lambda: ctx.ax(0, S.OP_HELMHOLTZ, u, au), 20, 3)
    t_dss = timed(lambda: ctx.dss(0, au), 20, 3)
    t_op = timed(lambda: ctx.opeval(0, S.OP_POISSON, S.X_VELOCITY, u, au),
                 20, 3)
    log(f"[bench] helmholtz {t_helm*1e6:.1f} us dss {t_dss*1e6:.1f} us "
        f"opeval {t_op*1e6:.1f} us")

    achieved = BYTES_PER_NODE_POISSON * nt / t_ax / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                "traffic": read_traffic(),
                "kernel": f"k_ax<{args.n},0> (FP64 Poisson Ax, launch "
                          f"variant {ctx.ax_variant(0, 0)[0]})",
                "bytes_per_launch": BYTES_PER_NODE_POISSON * nt}

    solves = {}
    if not args.no_solve:
        solves = solve_case(S, torch, ctx, mesh, args.max_iter)
    mg_orders = [ctx.level_n(l) - 1 for l in range(ctx.nlevels)]
    variants = {}
    for l in range(ctx.nlevels):
        v, us = ctx.ax_variant(l, S.PREC_FP64)
        variants[f"level{l}_n{ctx.level_n(l)}"] = {
            "variant": v, "us_per_variant": [round(x, 2) for x in us]}
    ctx.close()

    particles = None if args.no_particles else particle_case(S, torch,
                                                              stream)
    if particles:
        log(f"[bench] {particles['particles']} particles: "
            f"{particles['seconds']*1e3:.2f} ms per step, "
            f"{particles['max_passes']} passes; RK2 tracers "
            f"{particles['tracers_rk2']['seconds']*1e3:.2f} ms per step")
    batchres = batch_small = None
    if not args.no_batch:
        batchres = batch_case(S, torch)
        # many small problems, where one problem leaves most of the GPU idle
        batch_small = batch_case(S, torch, nprob=64, seed=12, mrange=(3, 6),
                                 points=(5, 6, 8))
    for br in (batchres, batch_small):
        if br:
            log(f"[bench] batch of {br['problems']} ({br['local_dofs']} "
                f"nodes): sequential {br['sequential_s']:.3f} s, batched "
                f"{br['batched_s']:.3f} s, with graphs "
                f"{br['batched_graph_s']:.3f} s")

    sweep = {}
    if not args.no_sweep:
        # the same problem at N = 5, 7, 11 on 48^3, 32^3, 20^3 elements
        for n, me in ((6, 48), (8, 32), (12, 20)):
            if n == args.n and me == m and solves:
                sweep[f"N{n-1}_{me}^3"] = solves
                continue
            sm = box_mesh(n, (me, me, me), deform=1, amp=0.1)
            sctx = S.Context(sm)
            sctx.set_stream(stream)
            sweep[f"N{n-1}_{me}^3"] = dict(
                local_dofs=sm.ntot,
                **solve_case(S, torch, sctx, sm, args.max_iter,
                             log_prefix=f"N={n-1} {me}^3 "))
            sctx.close()

    cpu = None if args.no_cpu else cpu_baseline(mesh, args.cpu_seconds)
    out = {
        "metric": "Ax GDOF/s + p-MG PCG time to 1e-10 residual, FP64 vs "
                  "mixed, 1 MI355X",
        "value": nt / t_ax / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_ax * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"deformed cube {m}^3 hexahedra, N={args.n-1}"
                               f" ({args.n}^3 GLL nodes/element), Dirichlet "
                               "Poisson, FP64 operator + FCG/p-MG solve",
                   "elements": m ** 3, "points_per_dir": args.n,
                   "local_dofs": nt, "unique_dofs": (m * (args.n - 1) + 1) ** 3,
                   "mg_orders": mg_orders},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "helmholtz_ax_gdofs": nt / t_helm / 1e9,
        "dss_us": t_dss * 1e6,
        "opeval_gdofs": nt / t_op / 1e9,
        "solve_fp64": solves.get("fp64"),
        "solve_mixed_fp32": solves.get("fp32"),
        "solve_mixed_bf16": solves.get("bf16"),
        "ax_launch_variants": variants,
        "particles": particles,
        "batch_uneven": batchres,
        "batch_small": batch_small,
        "mixed_precision_sweep": sweep or None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
