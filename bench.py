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
    reference position, reference-space velocity (3, 0, 0) (about the
    flow's: centreline speed 1 times 2/h for elements h ~ 0.5 long in x),
    response time tau ~ U[0.05, 0.5], advanced over one step dt = 0.05,
    so that a share of them changes element."""
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
    dt, T = 0.05, 1.0
    ip = lambda a: torch.from_numpy(a.astype(np.int32)).cuda()
    fp = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    el = rng.integers(0, mesh.nel, npart)
    xyz = rng.uniform(-1, 1, (3, npart))
    zero = np.zeros(npart)
    base = {"el": el, "xyz": xyz, "tau": rng.uniform(0.05, 0.5, npart)}

    def fresh():
        return {"el": ip(base["el"]), "old_el": ip(base["el"]),
                "x": fp(base["xyz"][0]), "y": fp(base["xyz"][1]),
                "z": fp(base["xyz"][2]), "u": fp(zero + 3.0), "v": fp(zero),
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


def channel_field(mesh, lx, lz, amp, seed):
    """Poiseuille u = 1 - y^2 plus a seeded divergence-free perturbation:
    (u', v') = amp (d phi/dy, -d phi/dx), w' = 0, with
    phi = (1 - y^2)^2 sum_m a_m sin(k_m x + s_m) cos(l_m z), k_m, l_m
    multiples of 2 pi / lx, 2 pi / lz (periodic), zero with its y-derivative
    on the walls y = +-1"""
    rng = np.random.default_rng(seed)
    X, Y, Z = (a.ravel() for a in (mesh.x, mesh.y, mesh.z))
    u = 1.0 - Y ** 2
    v = np.zeros_like(u)
    env, denv = (1 - Y ** 2) ** 2, -4 * Y * (1 - Y ** 2)
    for _ in range(4):
        k = 2 * np.pi / lx * rng.integers(1, 4)
        l = 2 * np.pi / lz * rng.integers(0, 3)
        a, ph = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
        sx, cx = np.sin(k * X + ph), np.cos(k * X + ph)
        u += amp * a * denv * sx * np.cos(l * Z) / 4
        v -= amp * a * env * k * cx * np.cos(l * Z) / 4
    return u, v, np.zeros_like(u)


def ns_tgv_case(S, torch, stream, steps=10):
    """The Taylor-Green configuration: periodic [0, 2 pi]^3 as 4^3
    hexahedra at N = 7 (n = 8, 32,768 nodes), u = sin x cos y cos z,
    v = -cos x sin y cos z, w = 0 plus seeded uniform noise of amplitude
    1e-3 (DSS-averaged), Re = 1600, dt = 1e-3, `steps` Navier-Stokes steps
    with the FCG + p-MG pressure solve (1e-10) and batched Helmholtz
    velocity solves.  Per-step wall time; the step's iteration counts."""
    from ssenss_amd.mesh import box_mesh, PERIODIC
    visc, dt = 1 / 1600, 1e-3
    mesh = box_mesh(8, (4, 4, 4), lo=(0, 0, 0), hi=(2 * np.pi,) * 3,
                    bc=(PERIODIC,) * 6, dt=dt, visc=visc)
    ctx = S.Context(mesh)
    ctx.set_stream(stream)
    X, Y, Z = (a.ravel() for a in (mesh.x, mesh.y, mesh.z))
    rng = np.random.default_rng(1600)
    levels = []
    for lev in range(3):
        d = np.exp(-3 * visc * (lev - 2) * dt)
        comps = []
        for c in (d * np.sin(X) * np.cos(Y) * np.cos(Z),
                  -d * np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X)):
            t = torch.from_numpy(c + 1e-3 * rng.uniform(-1, 1, c.shape)).cuda()
            ctx.dss_average(0, t)
            comps.append(t)
        levels.append(comps)
    res = {}
    with S.NavierStokes(ctx, dt, visc, 1.0, 0.0, (0, 0, 0), True) as ns:
        ns.init(levels)
        torch.cuda.synchronize()
        times, its = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            st = ns.step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            its.append([x.iterations for x in st])
        u = [ns.read(k) for k in range(3)]
        ke = 0.5 * sum(ctx.dot(0, c, c) for c in u) / ctx.dot(
            0, torch.ones_like(u[0]), torch.ones_like(u[0]))
        res = {"workload": "Taylor-Green, periodic 4^3 hexahedra, N=7, "
                           "Re=1600, dt=1e-3", "local_dofs": mesh.ntot,
               "steps": steps, "seconds_per_step_first": times[0],
               "seconds_per_step": float(np.median(times[1:])),
               "iterations_per_step": its[-1],
               "kinetic_energy_final": float(ke)}
    ctx.close()
    return res


def ns_channel_case(S, torch, stream, steps=3, ntracers=1_000_000,
                    prec="fp32", coarse_direct=0):
    """The channel configuration: 24x16x24 hexahedra at N = 9 (n = 10,
    9.2M nodes) on [0, 4 pi] x [-1, 1] x [0, 4 pi / 3], tanh wall
    stretching 2.0 in y, periodic in x and z, visc = 1/180 (Re_tau = 180
    with u_tau = h = 1), body force gx = 2 visc (the laminar profile's
    balance), dt = 2e-3; Poiseuille profile plus 10% divergence-free noise
    at all three history levels.  Each step: OIFS convection, FCG +
    p-MG pressure solve with the FP32 smoother (its coarsest level,
    9792 unique nodes, solved by the sparse direct factor, or by the dense
    inverse when coarse_direct >= that: SSX_OPT_COARSE_DIRECT), the three
    velocity Helmholtz solves as one batch, then the velocity history
    updated and `ntracers` RK2 tracers advanced over the step.  The first
    step is a warm-up (graph capture, reduced-precision copies, the coarse
    factor)."""
    from ssenss_amd.mesh import box_mesh, PERIODIC, NOSLIP
    visc, dt = 1 / 180, 2e-3
    lx, lz = 4 * np.pi, 4 * np.pi / 3
    mesh = box_mesh(10, (24, 16, 24), lo=(0.0, -1.0, 0.0), hi=(lx, 1.0, lz),
                    bc=(PERIODIC, PERIODIC, NOSLIP, NOSLIP, PERIODIC,
                        PERIODIC), deform=2, amp=2.0, dt=dt, visc=visc)
    t0 = time.perf_counter()
    ctx = S.Context(mesh)
    ctx.set_stream(stream)
    tset = time.perf_counter() - t0
    code = {"fp64": S.PREC_FP64, "fp32": S.PREC_FP32, "bf16": S.PREC_BF16}
    ctx.set_option(S.OPT_SMOOTHER_PRECISION, code[prec])
    ctx.set_option(S.OPT_COARSE_DIRECT, coarse_direct)
    u0 = channel_field(mesh, lx, lz, 0.1, 180)
    levels = []
    for _ in range(3):
        comps = []
        for c in u0:
            t = torch.from_numpy(np.ascontiguousarray(c)).cuda()
            ctx.dss_average(0, t)
            comps.append(t)
        levels.append(comps)
    rng = np.random.default_rng(181)
    ip = lambda a: torch.from_numpy(a.astype(np.int32)).cuda()
    fp = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    el = rng.integers(0, mesh.nel, ntracers)
    xyz = rng.uniform(-1, 1, (3, ntracers))
    zero = np.zeros(ntracers)
    tr = {"el": ip(el), "old_el": ip(el), "x": fp(xyz[0]), "y": fp(xyz[1]),
          "z": fp(xyz[2]), "u": fp(zero), "v": fp(zero), "w": fp(zero),
          "t": fp(zero), "reinit": ip(zero)}
    res = {}
    with S.NavierStokes(ctx, dt, visc, 1.0, 0.0, (2 * visc, 0, 0),
                        True) as ns:
        ns.init(levels)
        for lev in levels:                       # tracer velocity history
            ctx.particles_set_velocity(*lev)
        torch.cuda.synchronize()
        ts, tt, its, passes = [], [], [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            st = ns.step()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ctx.particles_set_velocity(*[ns.read(k) for k in range(3)])
            passes.append(ctx.tracers_step(tr, ns.time, dt))
            t2 = time.perf_counter()
            ts.append(t1 - t0)
            tt.append(t2 - t1)
            its.append([x.iterations for x in st])
        moved = int((tr["el"].cpu().numpy() != el).sum())
        res = {"workload": "channel 24x16x24 hexahedra, N=9, tanh 2.0, "
                           "Re_tau 180, dt 2e-3", "local_dofs": mesh.ntot,
               "context_setup_s": tset, "pressure_smoother": prec,
               "coarse_direct_max_unknowns": coarse_direct,
               "coarse_solve": ("dense inverse" if coarse_direct >= 9792
                                else "sparse factor"),
               "steps": steps, "seconds_per_step_first": ts[0],
               "seconds_per_step": float(np.median(ts[1:])),
               "iterations_per_step": its[-1],
               "tracers_rk2": {"tracers": ntracers,
                               "seconds_per_step": float(np.median(tt[1:])),
                               "max_passes": max(passes),
                               "changed_element_total": moved}}
    ctx.close()
    return res


def batch_mg_case(S, torch, nprob=64, seed=64, mrange=(8, 24),
                  points=(5, 6, 8, 10, 12), prec="fp32"):
    """The ensemble configuration: nprob independent Helmholtz problems on
    randomly deformed boxes (amplitude U(0, 0.15)), element counts m^3 with
    m uniform in mrange, orders N in {4, 5, 7, 9, 11} (n = N + 1), lambda ~
    U[1, 100], right-hand side the weak form of a seeded N(0, 1) nodal
    field, FCG + p-multigrid to 1e-10 relative with the FP64 outer loop and
    the reduced-precision smoother `prec`.  All problems at once
    (ssx_batch_solve: each V-cycle's level CGs of all problems in shared
    launches) against the same problems solved one by one (ssx_fcg); both
    after a warm-up solve of each problem."""
    from ssenss_amd.mesh import box_mesh
    code = {"fp64": S.PREC_FP64, "fp32": S.PREC_FP32, "bf16": S.PREC_BF16}
    rng = np.random.default_rng(seed)
    ctxs, rhs, shapes = [], [], []
    t0 = time.perf_counter()
    for _ in range(nprob):
        m = int(rng.integers(mrange[0], mrange[1] + 1))
        n = int(rng.choice(list(points)))
        lam = rng.uniform(1, 100)
        mesh = box_mesh(n, (m, m, m), deform=1, amp=rng.uniform(0, 0.15),
                        dt=(11.0 / 6.0) / lam, visc=1.0)
        c = S.Context(mesh)
        c.set_stream(torch.cuda.current_stream())
        c.set_option(S.OPT_SMOOTHER_PRECISION, code[prec])
        # the batch solves every coarsest level in one batched CG pass; the
        # one-by-one solves take the same coarse CG (not the sparse factor
        # a lone problem gets by default), so both run one algorithm
        c.set_option(S.OPT_COARSE_SPARSE, 0)
        # right-hand side: DSS(w detJ f) of a seeded N(0,1) nodal field,
        # drawn on the GPU (the mesh arrays stay on the host only for the
        # context set-up)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(rng.integers(1 << 30)))
        massv = torch.from_numpy(c.level_data(0, S.DATA_MASS)).cuda()
        b = torch.randn(c.ntot(0), dtype=torch.float64, device="cuda",
                        generator=gen) * massv
        c.dss(0, b)
        del mesh
        ctxs.append(c)
        rhs.append(b)
        shapes.append((m, n))
    tset = time.perf_counter() - t0
    log(f"[bench] ensemble of {nprob}: set-up {tset:.1f} s")
    eps = [1e-10 * float(np.sqrt(c.dot(0, b, b))) for c, b in zip(ctxs, rhs)]
    dofs = sum(c.ntot(0) for c in ctxs)
    t_seq, single_its = None, []
    for rep in range(2):                 # the first run is a warm-up
        single_its = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for c, b, e in zip(ctxs, rhs, eps):
            f = b.clone()
            u = torch.zeros_like(f)
            st = c.fcg(S.OP_HELMHOLTZ, S.X_VELOCITY, S.PC_MG, f, u, e, 500)
            single_its.append(st.iterations)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
    log(f"[bench] ensemble one by one {t_seq:.3f} s")
    batch = S.Batch([(c, S.OP_HELMHOLTZ, S.X_VELOCITY) for c in ctxs])
    t_b, b_its = None, []
    for rep in range(2):
        fs = [b.clone() for b in rhs]
        us = [torch.zeros_like(f) for f in fs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = batch.solve([S.PC_MG] * nprob, fs, us, eps, 500)
        t_b = time.perf_counter() - t0
        b_its = [x.iterations for x in st]
    log(f"[bench] ensemble batched {t_b:.3f} s")
    batch.close()
    for c in ctxs:
        c.close()
    return {"problems": nprob, "local_dofs": dofs,
            "elements_per_dir": list(mrange), "orders": sorted(
                {n - 1 for _, n in shapes}),
            "smoother_precision": prec, "context_setup_s": tset,
            "sequential_s": t_seq, "batched_s": t_b,
            "speedup_batched": t_seq / t_b,
            "iterations_single": single_its, "iterations_batched": b_its,
            "max_iteration_difference": max(
                abs(a - b) for a, b in zip(single_its, b_its))}


def reference_cpu(seconds):
    """The compiled reference (oracle/_ref/ref_harness, built from the
    SSENSS sources by oracle/Makefile) timed on the host: its Poisson
    element operator over all elements for about `seconds`, and one
    Dirichlet FCG + p-MG solve to 1e-10 relative, on a 8^3-element sample
    of the deformed cube at n = 8, one core (MKL sequential, one rank);
    None where the harness or its libraries are absent."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(exe, os.X_OK):
        return None
    from ssenss_amd.mesh import box_mesh
    from oracle import pyoracle as P
    sm = box_mesh(8, (8, 8, 8), deform=1, amp=0.1)
    f = smooth_source(sm.x.ravel(), sm.y.ravel(), sm.z.ravel())
    o = P.Oracle(sm)
    _, rho, _, _ = P.gll_basis(8)
    w3 = (rho[None, None, :] * rho[None, :, None] * rho[:, None, None]).ravel()
    b = o.dss(0, np.tile(w3, sm.nel) * o.data(0, P.DETJ) * f)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "mesh.bin"), "wb") as fh:
            fh.write(np.array([sm.n, sm.nel], np.int32).tobytes())
            fh.write(np.array([1.0, 1.0], np.float64).tobytes())
            fh.write(np.array([(2 * sm.n) // 3], np.int32).tobytes())
            fh.write(np.array([0.0], np.float64).tobytes())
            for a in (sm.neigh, sm.edge_neigh, sm.edge_size, sm.corn_neigh,
                      sm.corn_size):
                fh.write(np.ascontiguousarray(a, np.int32).tobytes())
            for a in (sm.x, sm.y, sm.z):
                fh.write(np.ascontiguousarray(a, np.float64).tobytes())
        b.tofile(os.path.join(d, "in_f.bin"))
        env = dict(os.environ, MKL_NUM_THREADS="1",
                   MKL_THREADING_LAYER="SEQUENTIAL", OMP_NUM_THREADS="1")
        try:
            r = subprocess.run([exe, "time", d, str(seconds)], env=env,
                               capture_output=True, text=True, timeout=300)
        except (OSError, subprocess.SubprocessError) as e:
            log(f"[bench] reference harness did not run: {e}")
            return None
    kv = r.stdout.split()
    if r.returncode != 0 or "ax_seconds" not in kv:
        log(f"[bench] reference harness failed: {r.stderr[-300:]}")
        return None
    vals = {kv[i]: float(kv[i + 1]) for i in range(0, len(kv) - 1, 2)}
    # the port on the same sample, one thread and all threads
    port = {}
    for th in (1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()
                                     or 1))):
        P.set_threads(th)
        o.set_diag_choice(0)
        eps = 1e-10 * np.sqrt(o.dot(0, b, b))
        t0 = time.perf_counter()
        _, it, _, _ = o.fcg(P.OP_POISSON, P.X_VELOCITY, P.PC_MG, b,
                            np.zeros(sm.ntot), eps, 500)
        port[P.get_threads()] = (time.perf_counter() - t0, it)
    o.close()
    return {"local_dofs": sm.ntot, "ax_seconds": vals["ax_seconds"],
            "ax_reps": int(vals["ax_reps"]),
            "solve_seconds": vals["solve_seconds"],
            "port_solve_seconds": {str(k): v[0] for k, v in port.items()},
            "port_iterations": {str(k): v[1] for k, v in port.items()}}


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
    ap.add_argument("--no-ns", action="store_true",
                    help="skip the Navier-Stokes workloads")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-problem p-MG ensemble")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the N = 5 / 7 / 11 mixed-precision sweep")
    ap.add_argument("--profile", action="store_true",
                    help="kernels only, for rocprofv3 counter runs: the "
                    "Poisson Ax on the workload, then its DSS, two outer "
                    "FCG iterations with FP64 and FP32 smoothers, then a 1 GiB "
                    "device copy (known bytes, to calibrate FETCH_SIZE/"
                    "WRITE_SIZE)")
    ap.add_argument("--profile-case", default="cube",
                    choices=("cube", "solve32", "solve32_bf16", "particles",
                             "ns"),
                    help="with --profile: the workload whose kernels run "
                    "(cube: the solver kernels above; solve32: one whole "
                    "FCG + p-MG solve of the cube with the FP32 smoother; "
                    "solve32_bf16: the same with BF16 operator factors; "
                    "particles: the 1M-particle channel step and RK2 "
                    "tracers; ns: two channel Navier-Stokes steps), each "
                    "followed by the 1 GiB calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this "
              "engine runs all elements on a single MI355X and has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import ssenss_amd as S
    from ssenss_amd.mesh import box_mesh

    def calibration_copy():
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)                                   # 1 GiB each way
        torch.cuda.synchronize()

    if args.profile and args.profile_case not in ("cube", "solve32",
                                                  "solve32_bf16"):
        stream = torch.cuda.current_stream()
        if args.profile_case == "particles":
            r = particle_case(S, torch, stream)
            log(f"[bench] particles {r['seconds']*1e3:.2f} ms, tracers "
                f"{r['tracers_rk2']['seconds']*1e3:.2f} ms")
        else:
            r = ns_channel_case(S, torch, stream, steps=2)
            log(f"[bench] NS channel {r['seconds_per_step']:.3f} s per step")
        calibration_copy()
        log("[bench] profile mode done")
        return

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

    warm_log = {}

    def timed(fn, steps, warmup, name=None, min_warm_s=0.1):
        """K timed launches after at least `warmup` untimed ones, and after
        at least min_warm_s of back-to-back launches: the first tens of
        milliseconds of GPU work after an idle spell run slower (about
        195 against 160 us per 32^3 n=8 operator launch, measured launch
        by launch) while the device's clocks settle, and a 20-launch window
        right after set-up measures that transient, not the kernel."""
        done = 0
        t0 = time.perf_counter()
        while True:
            for _ in range(max(warmup, 1)):
                fn()
            done += max(warmup, 1)
            torch.cuda.synchronize()
            if done >= warmup and time.perf_counter() - t0 >= min_warm_s:
                break
        if name:
            warm_log[name] = done
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    if args.profile and args.profile_case.startswith("solve32"):
        prec = "bf16" if args.profile_case.endswith("bf16") else "fp32"
        solve_case(S, torch, ctx, mesh, args.max_iter, precs=(prec,),
                   log_prefix="profile ")
        calibration_copy()
        log("[bench] profile mode done")
        ctx.close()
        return
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
        calibration_copy()
        log("[bench] profile mode done")
        ctx.close()
        return
    t_ax = timed(lambda: ctx.ax(0, S.OP_POISSON, u, au), args.steps,
                 args.warmup, "ax")
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
    ns_tgv = ns_chan = None
    if not args.no_ns:
        ns_tgv = ns_tgv_case(S, torch, stream)
        log(f"[bench] NS Taylor-Green: {ns_tgv['seconds_per_step']*1e3:.1f} "
            f"ms per step")
        ns_chan = ns_channel_case(S, torch, stream)
        log(f"[bench] NS channel: {ns_chan['seconds_per_step']:.3f} s per "
            f"step, tracers {ns_chan['tracers_rk2']['seconds_per_step']*1e3:.2f}"
            f" ms")
    ensemble = None if args.no_ensemble else batch_mg_case(S, torch)
    batch_small = None
    if not args.no_batch:
        # many small problems, where one problem leaves most of the GPU idle
        batch_small = batch_case(S, torch, nprob=64, seed=12, mrange=(3, 6),
                                 points=(5, 6, 8))
    for br in (batch_small,):
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

    cpu, cpu_ttr = None, None
    if not args.no_cpu:
        ref = reference_cpu(3.0)
        if ref:
            cpu = {"value": ref["local_dofs"] / ref["ax_seconds"] / 1e9,
                   "unit": "GDOF/s", "cores": 1, "kind": "reference",
                   "sample": f"the compiled reference's Poisson element "
                             f"operator over a 8^3-element deformed cube "
                             f"at n=8 ({ref['local_dofs']} local nodes), "
                             f"{ref['ax_reps']} applications, one core"}
            gpu = solves.get("fp64")
            cpu_ttr = {
                "sample": "Dirichlet Poisson FCG + p-MG to 1e-10 relative, "
                          "8^3 deformed cube, n=8",
                "local_dofs": ref["local_dofs"],
                "reference_1core_s": ref["solve_seconds"],
                "port_s_by_threads": ref["port_solve_seconds"],
                "port_iterations": ref["port_iterations"],
                "reference_ns_per_dof": ref["solve_seconds"]
                / ref["local_dofs"] * 1e9,
                "gpu_fp64_ns_per_dof": (gpu["time_to_residual_s"] / nt * 1e9
                                        if gpu else None)}
        else:
            cpu = cpu_baseline(mesh, args.cpu_seconds)
    out = {
        "metric": "Ax GDOF/s + p-MG PCG time to 1e-10 residual, FP64 vs "
                  "mixed, 1 MI355X",
        "value": nt / t_ax / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "warmup_launches_ax": warm_log.get("ax"),
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
        "cpu_time_to_residual": cpu_ttr,
        "helmholtz_ax_gdofs": nt / t_helm / 1e9,
        "dss_us": t_dss * 1e6,
        "opeval_gdofs": nt / t_op / 1e9,
        "solve_fp64": solves.get("fp64"),
        "solve_mixed_fp32": solves.get("fp32"),
        "solve_mixed_bf16": solves.get("bf16"),
        "ax_launch_variants": variants,
        "particles": particles,
        "batch_mg_ensemble": ensemble,
        "batch_small": batch_small,
        "ns_taylor_green": ns_tgv,
        "ns_channel": ns_chan,
        "mixed_precision_sweep": sweep or None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
