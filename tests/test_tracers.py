"""RK2 tracers (ssx_tracers_step; oracle orc_tracers_step).

The tracer mode has no counterpart in the reference (its particles are
inertial, RK4), so it is pinned by closed forms on the oracle:
- uniform flow: every tracer moves by U*T exactly, across any number of
  elements (straight moves to the boundary and hops are exact);
- a linear field u = B xi inside one element: one step is Heun's
  xi + h B xi + h^2/2 B^2 xi exactly (the spectral interpolation of a
  linear field is exact);
- a tracer pushed into a wall ends on the wall, in its element.
The GPU kernel is then held to the oracle: owning elements exact, states to
rounding (the engine's GLL basis and J^-1 agree with the oracle's to ~1e-15,
amplified by the walk's straight moves on a few trajectories)."""
import ctypes as C
import numpy as np
import pytest

import ssenss_amd as S
from ssenss_amd.mesh import box_mesh, NOSLIP, PERIODIC
from oracle import pyoracle as P


def _records(el, xyz, t0):
    n = len(el)
    rec = np.zeros((n, 9))
    rec[:, 0] = el
    rec[:, 1:4] = xyz
    rec[:, 7] = t0
    rec[:, 8] = 1.0
    return rec


def _uniform_case(n):
    L = (1.0, 0.75, 1.25)
    mesh = box_mesh(n, (4, 3, 5), lo=(0, 0, 0), hi=L, bc=(PERIODIC,) * 6)
    U = np.array([0.37, -0.21, 0.53])
    nt = mesh.ntot
    vel = np.concatenate([np.full(nt, U[0]), np.full(nt, U[1]),
                          np.full(nt, U[2])])
    rng = np.random.default_rng(5 + n)
    m = 400
    rec = _records(rng.integers(0, mesh.nel, m),
                   rng.uniform(-0.999, 0.999, (m, 3)), 0.0)
    return mesh, np.array(L), U, vel, rec


def _periodic_diff(a, b, L):
    return (a - b + 0.5 * L) % L - 0.5 * L


@pytest.mark.parametrize("n", [4, 8])
def test_tracers_uniform_flow_closed_form(n):
    mesh, L, U, vel, rec = _uniform_case(n)
    o = P.Oracle(mesh)
    x0 = o.positions(rec)
    steps, dt = 8, 0.11
    out, re, passes, xyz = o.tracers(vel, (1.0, 1.0, 1.0), rec, steps * dt,
                                     dt, steps=steps)
    assert re.sum() == 0 and passes >= 2
    assert (out[:, 0] != rec[:, 0]).sum() > 300          # crossings happen
    d = _periodic_diff(xyz, x0 + U * steps * dt, L)
    assert np.abs(d).max() <= 1e-12
    assert np.allclose(out[:, 7], steps * dt, rtol=0, atol=0)
    # u, v, w: the fluid velocity in reference coordinates, (2/h) U
    h = L / np.array([4, 3, 5])
    assert np.abs(out[:, 4:7] - 2.0 * U / h).max() <= 1e-12


@pytest.mark.parametrize("n", [3, 8])
def test_tracers_linear_field_heun_closed_form(n):
    mesh = box_mesh(n, (1, 1, 1), lo=(0, 0, 0), hi=(1, 1, 1))
    w = 0.8
    B = np.array([[0.0, -w, 0.1], [w, 0.0, -0.2], [0.05, 0.3, 0.0]])
    c = np.array([0.5, 0.5, 0.5])
    X = np.stack([mesh.x.ravel(), mesh.y.ravel(), mesh.z.ravel()], 1)
    # xi = 2 (x - c) and J^-1 = 2 I: u_phys = B (x - c) gives dxi/dt = B xi
    up = (X - c) @ B.T
    vel = np.concatenate([up[:, 0], up[:, 1], up[:, 2]])
    rng = np.random.default_rng(9)
    m = 200
    xi0 = rng.uniform(-0.5, 0.5, (m, 3))
    h = 0.05
    rec = _records(np.zeros(m), xi0, 1.0 - h)
    o = P.Oracle(mesh)
    out, re, passes, _ = o.tracers(vel, (1.0, 1.0, 1.0), rec, 1.0, h)
    ref = xi0 + h * xi0 @ B.T + 0.5 * h * h * xi0 @ (B @ B).T
    assert re.sum() == 0 and passes == 1
    assert np.abs(out[:, 1:4] - ref).max() <= 1e-13
    assert np.abs(out[:, 4:7] - ref @ B.T).max() <= 1e-13


def test_tracers_stop_at_a_wall():
    mesh = box_mesh(6, (2, 1, 1), lo=(0, 0, 0), hi=(2, 1, 1))
    nt = mesh.ntot
    vel = np.concatenate([np.full(nt, 1.0), np.zeros(nt), np.zeros(nt)])
    rec = _records(np.array([1, 1, 0]), np.array([[0.9, 0.2, -0.3],
                                                  [-0.9, 0.0, 0.0],
                                                  [0.5, 0.1, 0.1]]), 0.0)
    o = P.Oracle(mesh)
    out, re, passes, xyz = o.tracers(vel, (1.0, 1.0, 1.0), rec, 0.5, 0.5)
    # el 1 at xi = 0.9 would pass x = 2 (the wall): stops there
    assert out[0, 0] == 1 and out[0, 1] == 1.0 and out[0, 7] == 0.5
    assert np.abs(xyz[0] - [2.0, 0.6, 0.35]).max() <= 1e-14
    # the others move by 0.5 in x; el 0 at xi 0.5 (x = 0.75) crosses into 1
    assert np.abs(xyz[1] - [1.55, 0.5, 0.5]).max() <= 1e-14
    assert out[2, 0] == 1 and np.abs(xyz[2] - [1.25, 0.55, 0.55]).max() <= 1e-14
    assert re.sum() == 0


# ---------------------------------------------------------------- GPU ---

def _gpu_tracers(ctx, vel, scales, rec, t_now, dt, steps=1):
    import torch
    nt = ctx.ntot(0)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    ip = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()
    for s in scales:
        ctx.particles_set_velocity(dev(s * vel[:nt]), dev(s * vel[nt:2 * nt]),
                                   dev(s * vel[2 * nt:]))
    p = {"el": ip(rec[:, 0]), "old_el": ip(rec[:, 0]), "x": dev(rec[:, 1]),
         "y": dev(rec[:, 2]), "z": dev(rec[:, 3]), "u": dev(rec[:, 4]),
         "v": dev(rec[:, 5]), "w": dev(rec[:, 6]), "t": dev(rec[:, 7]),
         "reinit": ip(np.zeros(len(rec)))}
    passes = 0
    for k in range(steps):
        passes = max(passes, ctx.tracers_step(p, t_now - (steps - 1 - k) * dt,
                                              dt))
    torch.cuda.synchronize()
    out = np.stack([p["el"].cpu().numpy().astype(np.float64)] + [
        p[k].cpu().numpy() for k in ("x", "y", "z", "u", "v", "w", "t")], 1)
    return out, p["reinit"].cpu().numpy(), passes


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4, 8])
def test_gpu_tracers_uniform_flow_closed_form(n):
    mesh, L, U, vel, rec = _uniform_case(n)
    o = P.Oracle(mesh)
    x0 = o.positions(rec)
    steps, dt = 8, 0.11
    with S.Context(mesh) as ctx:
        out, re, passes = _gpu_tracers(ctx, vel, (1.0, 1.0, 1.0), rec,
                                       steps * dt, dt, steps)
    d = _periodic_diff(o.positions(out), x0 + U * steps * dt, L)
    print(f"[gpu] tracers uniform n={n}: passes {passes}, max err "
          f"{np.abs(d).max():.1e}")
    assert re.sum() == 0 and np.abs(d).max() <= 1e-12


@pytest.mark.gpu
def test_gpu_tracers_against_oracle():
    """20k tracers in a Taylor-Green field on a periodic 4^3 mesh at n = 8,
    3 steps with several hops: elements exact, states to rounding; the
    host-array entry gives the device entry's result bit for bit"""
    import torch
    mesh = box_mesh(8, (4, 4, 4), lo=(0, 0, 0), hi=(2 * np.pi,) * 3,
                    bc=(PERIODIC,) * 6)
    X, Y, Zc = mesh.x, mesh.y, mesh.z
    vel = np.concatenate([(2 * np.sin(X) * np.cos(Y) * np.cos(Zc)).ravel(),
                          (-2 * np.cos(X) * np.sin(Y) * np.cos(Zc)).ravel(),
                          (np.sin(X + Zc)).ravel()])
    rng = np.random.default_rng(41)
    n = 20000
    rec = _records(rng.integers(0, mesh.nel, n),
                   rng.uniform(-0.99, 0.99, (n, 3)), 1.0 - 3 * 0.25)
    scales = (0.8, 0.9, 1.0)
    o = P.Oracle(mesh)
    ref, re_o, passes_o, _ = o.tracers(vel, scales, rec, 1.0, 0.25, steps=3)
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        out, re, passes = _gpu_tracers(ctx, vel, scales, rec, 1.0, 0.25, 3)
        # host-array entry, same velocity history
        hp = {"el": rec[:, 0].astype(np.int32), "old_el":
              rec[:, 0].astype(np.int32), "reinit": np.zeros(n, np.int32)}
        for k, col in zip(("x", "y", "z", "u", "v", "w", "t"), range(1, 8)):
            hp[k] = np.ascontiguousarray(rec[:, col])
        nt = ctx.ntot(0)
        for s in scales:
            ctx.particles_set_velocity_host(s * vel[:nt], s * vel[nt:2 * nt],
                                            s * vel[2 * nt:])
        for k in range(3):
            ctx.tracers_step_host(hp, 1.0 - (2 - k) * 0.25, 0.25)
    out_h = np.stack([hp["el"].astype(np.float64)] + [
        hp[k] for k in ("x", "y", "z", "u", "v", "w", "t")], 1)
    hops = (out[:, 0] != rec[:, 0]).sum()
    d = np.abs(out[:, 1:] - ref[:, 1:]).max(axis=1)
    print(f"[gpu] 20k tracers: {hops} changed element, passes {passes} "
          f"(oracle {passes_o}), state diff median {np.median(d):.1e} "
          f"max {d.max():.1e}")
    assert hops > 5000
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert np.array_equal(re, re_o) and re.sum() == 0
    assert np.median(d) <= 1e-12 and d.max() <= 1e-9
    assert np.array_equal(out, out_h)
