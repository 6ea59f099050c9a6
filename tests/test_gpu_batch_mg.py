"""Batched solves with the p-multigrid V-cycle (ssx_batch_solve) through the
C ABI: flexible CG + p-MG problems advancing together, their smoother and
coarsest-level CGs in shared launches, Jacobi-PCG problems riding along.

Bars: each multigrid problem's FCG iteration count within one of the
reference's (golden meshes) and of the single-problem ssx_fcg on the same
context; solutions within the solve tolerance (1e-8 relative to the
reference's or the single solve's, both solved to 1e-10 ||b||_w); the
Jacobi problems' iteration counts within one of ssx_cg."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from ssenss_amd.mesh import box_mesh  # noqa: E402
from conftest import golden, golden_mesh, rel, CASES  # noqa: E402
from oracle import pyoracle as P  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_ctx(mesh):
    ctx = S.Context(mesh)
    ctx.set_stream(torch.cuda.current_stream())
    return ctx


def _batch_solve(probs, precon, rhs, eps, max_iter=5000):
    batch = S.Batch(probs)
    try:
        fs = [dev(r) for r in rhs]
        us = [torch.zeros_like(f) for f in fs]
        st = batch.solve(precon, fs, us, eps, max_iter)
        return st, [host(u) for u in us]
    finally:
        batch.close()


@pytest.mark.parametrize("var", [S.PRESSURE, S.X_VELOCITY])
def test_batch_mg_golden_against_reference(var):
    """the five golden meshes (orders 5, 6, 8, 12; Dirichlet, periodic and
    channel) as one batch of FCG + p-MG Poisson solves, pressure (null space
    removed) or Dirichlet (masked, Poisson diagonal): iterations within one
    of the reference's, solutions within 1e-8 of its"""
    cases = [(nm, golden(nm)) for nm in CASES]
    ctxs = [make_ctx(golden_mesh(Z)) for _, Z in cases]
    try:
        for c in ctxs:
            c.set_diag_choice(-1 if var == S.PRESSURE else 0)
        probs = [(c, S.OP_POISSON, var) for c in ctxs]
        rhs = [Z["in_f"] for _, Z in cases]
        eps = [Z["ref_scalars"][0] for _, Z in cases]
        print(f"[gpu] batch p-MG golden var {var}: {len(probs)} problems",
              flush=True)
        st, sol = _batch_solve(probs, [S.PC_MG] * len(probs), rhs, eps)
        key, col = (("ref_fcg_pressure", 3) if var == S.PRESSURE
                    else ("ref_fcg_dirichlet", 5))
        for k, (nm, Z) in enumerate(cases):
            it_ref = int(Z["ref_scalars"][col])
            err = rel(sol[k], Z[key])
            print(f"[gpu]   {nm}: {st[k].iterations} its (reference "
                  f"{it_ref}), vcycles {st[k].vcycles}, coarse "
                  f"{st[k].coarse_iters}, |u-u_ref|/|u_ref| {err:.1e}")
            assert st[k].converged and st[k].residual <= eps[k]
            assert abs(st[k].iterations - it_ref) <= 1
            assert err <= 1e-8
    finally:
        for c in ctxs:
            c.set_diag_choice(-1)
            c.close()


def _random_problems(rng, count, orders, sizes):
    meshes, lams = [], []
    for _ in range(count):
        n = int(rng.choice(orders))
        m = int(rng.choice(sizes))
        lam = rng.uniform(1, 100)
        meshes.append(box_mesh(n, (m, m, m), deform=1,
                               amp=rng.uniform(0, 0.15),
                               dt=(11.0 / 6.0) / lam, visc=1.0))
        lams.append(lam)
    return meshes, lams


def _assembled_rhs(ctx, rng):
    b = dev(rng.standard_normal(ctx.ntot(0)) * ctx.level_data(0, S.DATA_MASS))
    ctx.dss(0, b)
    return host(b)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_batch_mg_uneven_helmholtz_matches_single(prec):
    """Helmholtz problems of orders 4, 5, 7, 9, 11 (n = 5..12) on randomly
    deformed boxes of uneven element counts and lambda in [1, 100], FCG +
    p-MG with the smoother in FP64 (0), FP32 (1) or FP32 with BF16 factors
    (2): each problem's iterations within one of ssx_fcg on its context with
    the same smoother, its solution within the solve tolerance of that
    solve's"""
    rng = np.random.default_rng(44 + prec)
    meshes, _ = _random_problems(rng, 7, [5, 6, 8, 10, 12], [2, 3, 4])
    # every order at least once
    for k, n in enumerate([5, 6, 8, 10, 12]):
        m = meshes[k]
        if m.n != n:
            meshes[k] = box_mesh(n, (2, 3, 2), deform=1, amp=0.1,
                                 dt=m.dt, visc=1.0)
    ctxs = [make_ctx(m) for m in meshes]
    try:
        for c in ctxs:
            c.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        rhs = [_assembled_rhs(c, rng) for c in ctxs]
        eps = [1e-10 * np.sqrt(c.dot(0, dev(r), dev(r)))
               for c, r in zip(ctxs, rhs)]
        probs = [(c, S.OP_HELMHOLTZ, S.X_VELOCITY) for c in ctxs]
        print(f"[gpu] batch p-MG Helmholtz prec {prec}: orders "
              f"{[m.n for m in meshes]}, elements "
              f"{[m.nel for m in meshes]}", flush=True)
        st, sol = _batch_solve(probs, [S.PC_MG] * len(probs), rhs, eps)
        for k, c in enumerate(ctxs):
            f = dev(rhs[k])
            u = torch.zeros_like(f)
            s1 = c.fcg(S.OP_HELMHOLTZ, S.X_VELOCITY, S.PC_MG, f, u, eps[k],
                       5000)
            err = rel(sol[k], host(u))
            print(f"[gpu]   n={meshes[k].n} nel={meshes[k].nel}: "
                  f"{st[k].iterations} its (single {s1.iterations}), "
                  f"|u-u1|/|u1| {err:.1e}")
            assert st[k].converged and s1.converged
            assert abs(st[k].iterations - s1.iterations) <= 1
            assert err <= 1e-8
    finally:
        for c in ctxs:
            c.set_option(S.OPT_SMOOTHER_PRECISION, 0)
            c.close()


def test_batch_step_pressure_and_velocities():
    """the solves of one Navier-Stokes step in one batch: the pressure
    Poisson FCG + p-MG (navier_stokes.c:157) and the three velocity
    Helmholtz Jacobi-PCG solves (navier_stokes.c:219) on one context; each
    against its single-problem solve (iterations within one, solutions
    within 1e-8)"""
    Z = golden("channel_n6")
    rng = np.random.default_rng(9)
    with make_ctx(golden_mesh(Z)) as ctx:
        rhs = [Z["in_f"]] + [_assembled_rhs(ctx, rng) for _ in range(3)]
        eps = [1e-10 * np.linalg.norm(r) for r in rhs]
        probs = [(ctx, S.OP_POISSON, S.PRESSURE)] + [
            (ctx, S.OP_HELMHOLTZ, v) for v in (S.X_VELOCITY, S.Y_VELOCITY,
                                               S.Z_VELOCITY)]
        precon = [S.PC_MG, S.PC_JACOBI, S.PC_JACOBI, S.PC_JACOBI]
        st, sol = _batch_solve(probs, precon, rhs, eps)
        for k, (c, op, var) in enumerate(probs):
            f = dev(rhs[k])
            u = torch.zeros_like(f)
            if precon[k] == S.PC_MG:
                s1 = c.fcg(op, var, S.PC_MG, f, u, eps[k], 5000)
            else:
                s1 = c.cg(0, op, var, f, u, eps[k], 5000)
            err = rel(sol[k], host(u))
            print(f"[gpu] step batch problem {k}: {st[k].iterations} its "
                  f"(single {s1.iterations}), |u-u1|/|u1| {err:.1e}")
            assert st[k].converged
            assert abs(st[k].iterations - s1.iterations) <= 1
            assert err <= 1e-8


def test_fcg_mg_helmholtz_against_oracle():
    """ssx_fcg(OP_HELMHOLTZ, PC_MG) -- p-multigrid preconditioned FCG for the
    Helmholtz operator, an extension of the reference's Jacobi-PCG velocity
    solve -- against the oracle's FCG + p-MG on the same deformed mesh:
    iterations within one, solutions within 1e-8"""
    mesh = box_mesh(6, (3, 2, 3), deform=1, amp=0.12, dt=(11.0 / 6.0) / 37.0,
                    visc=1.0)
    o = P.Oracle(mesh)
    rng = np.random.default_rng(3)
    b = o.dss(0, rng.standard_normal(mesh.ntot))
    eps = 1e-10 * np.sqrt(o.dot(0, b, b))
    u_o, it_o, _, _ = o.fcg(P.OP_HELMHOLTZ, P.X_VELOCITY, P.PC_MG, b.copy(),
                            np.zeros(mesh.ntot), eps, 500)
    with make_ctx(mesh) as ctx:
        f = dev(b)
        u = torch.zeros_like(f)
        st = ctx.fcg(S.OP_HELMHOLTZ, S.X_VELOCITY, S.PC_MG, f, u, eps, 500)
        err = rel(host(u), u_o)
    print(f"[gpu] fcg p-MG Helmholtz: {st.iterations} its (oracle {it_o}), "
          f"|u-u_o|/|u_o| {err:.1e}")
    assert st.converged
    assert abs(st.iterations - it_o) <= 1
    assert err <= 1e-8
