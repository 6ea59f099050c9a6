"""The particle and tracer walks in all their forms give the same bits: one
relocation pass per launch with the moving particles compacted between
passes (SSX_OPT_PARTICLE_WALK 1, the default), the whole walk in one launch
in the sorted order (0), and the whole walk one lane per particle as stored
(the no-allocation host entry).  Fast, heavy particles on a periodic mesh,
so that many of them change element and some take every pass; the full
tensor-product interpolation (n = 4) and the 5-point stencil (n = 8)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from ssenss_amd.mesh import box_mesh, PERIODIC  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS_F = ("x", "y", "z", "u", "v", "w", "t", "tau")


def _case(n, npart, seed):
    mesh = box_mesh(n, (4, 4, 4), lo=(0, 0, 0), hi=(2 * np.pi,) * 3,
                    bc=(PERIODIC,) * 6, deform=1, amp=0.1)
    X, Y, Z = mesh.x, mesh.y, mesh.z
    vel = [(np.sin(X) * np.cos(Y) * np.cos(Z)).ravel(),
           (-np.cos(X) * np.sin(Y) * np.cos(Z)).ravel(),
           (0.5 * np.sin(X + Z)).ravel()]
    rng = np.random.default_rng(seed)
    # elements drawn unevenly (some empty), a few particles sharing a spot
    el = (rng.integers(0, mesh.nel // 2, npart) * 2).astype(np.int32)
    rec = {"el": el, "old_el": el.copy(),
           "reinit": np.zeros(npart, np.int32)}
    xyz = rng.uniform(-0.99, 0.99, (3, npart))
    xyz[:, :8] = 0.25
    for k, a in zip(("x", "y", "z"), xyz):
        rec[k] = np.ascontiguousarray(a)
    uvw = 4.0 * rng.standard_normal((3, npart))
    for k, a in zip(("u", "v", "w"), uvw):
        rec[k] = np.ascontiguousarray(a)
    rec["t"] = np.full(npart, 1.0 - 0.2)
    rec["tau"] = np.exp(rng.uniform(np.log(0.1), 0.0, npart))
    return mesh, vel, rec


def _set_velocity(ctx, vel):
    for s in (0.8, 0.9, 1.0):
        ctx.particles_set_velocity(*[torch.from_numpy(s * v).cuda()
                                     for v in vel])


def _on_device(rec):
    return {k: torch.from_numpy(v.copy()).cuda() for k, v in rec.items()}


def _to_host(p):
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in p.items()}


def _same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("n", [4, 8])
def test_particle_walk_orders_bit_identical(n):
    mesh, vel, rec = _case(n, 30000, 5 + n)
    out = {}
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        _set_velocity(ctx, vel)
        for mode in (1, 0):
            ctx.set_option(S.OPT_PARTICLE_WALK, mode)
            p = _on_device(rec)
            passes = ctx.particles_step(p, 1.0, 0.2)
            out[mode] = (_to_host(p), passes)
        h = {k: v.copy() for k, v in rec.items()}
        passes_h = ctx.particles_step_host_noalloc(h, 1.0, 0.2)
    hops = int((out[1][0]["el"] != rec["el"]).sum())
    print(f"[gpu] particle walk n={n}: {hops} of {len(rec['el'])} changed "
          f"element, passes {out[1][1]} / {out[0][1]} / {passes_h}")
    assert hops > 3000
    assert out[1][1] == out[0][1] == passes_h
    assert _same(out[1][0], out[0][0])
    assert _same(out[1][0], h)


@pytest.mark.parametrize("n", [4, 8])
def test_tracer_walk_orders_bit_identical(n):
    mesh, vel, rec = _case(n, 30000, 17 + n)
    out = {}
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        _set_velocity(ctx, vel)
        for mode in (1, 0):
            ctx.set_option(S.OPT_PARTICLE_WALK, mode)
            p = _on_device(rec)
            passes = ctx.tracers_step(p, 1.0, 0.2)
            out[mode] = (_to_host(p), passes)
    hops = int((out[1][0]["el"] != rec["el"]).sum())
    print(f"[gpu] tracer walk n={n}: {hops} changed element, passes "
          f"{out[1][1]} / {out[0][1]}")
    assert hops > 1000
    assert out[1][1] == out[0][1]
    assert _same(out[1][0], out[0][0])


def test_particle_walk_counts_bad_elements_once():
    """particles naming no element land in the sort's last bucket: every
    form of the walk reports each once and leaves them untouched"""
    mesh, vel, rec = _case(8, 4000, 3)
    rec["el"][[7, 100, 2500]] = [-1, mesh.nel, 10 ** 6]
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        _set_velocity(ctx, vel)
        for mode in (1, 0):
            ctx.set_option(S.OPT_PARTICLE_WALK, mode)
            p = _on_device(rec)
            with pytest.raises(S.SsxError, match="3 particles"):
                ctx.particles_step(p, 1.0, 0.2)
            h = _to_host(p)
            assert h["el"][100] == mesh.nel and h["x"][100] == rec["x"][100]
