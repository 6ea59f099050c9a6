"""ssx_particles_step_host_noalloc (staging in a fixed device array, no
device allocation) against ssx_particles_step_host: bit-identical particle
states, owning elements and pass counts on two calls in a row, the second
with fewer particles; sizes beyond the stated maximum and a context whose
particle tables are not set up yet are refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from conftest import golden, golden_mesh  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = ("x", "y", "z", "u", "v", "w", "t", "tau")


def _ctx(mesh):
    ctx = S.Context(mesh)
    ctx.set_stream(torch.cuda.current_stream())
    return ctx


def _particles(rec):
    p = {"el": np.ascontiguousarray(rec[:, 0], np.int32),
         "old_el": np.ascontiguousarray(rec[:, 0], np.int32),
         "reinit": np.zeros(len(rec), np.int32)}
    for c, k in enumerate(COLS):
        p[k] = np.ascontiguousarray(rec[:, 1 + c], np.float64)
    return p


def _head(p, m):
    return {k: np.ascontiguousarray(v[:m]) for k, v in p.items()}


def _same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


def _set_velocity(ctx, Z):
    h, vel = Z["particle_hdr"], Z["in_vel"]
    nt = ctx.ntot(0)
    for s in h[2:]:
        ctx.particles_set_velocity_host(s * vel[:nt], s * vel[nt:2 * nt],
                                        s * vel[2 * nt:])


def test_noalloc_particle_step_matches_allocating_entry():
    Z = golden("channel_n6")
    h, rec = Z["particle_hdr"], Z["in_particles"]
    t1, dt = float(h[0]), float(h[1])
    t2 = t1 + dt
    m2 = len(rec) // 3
    print(f"[gpu] no-alloc particle step: channel_n6, {len(rec)} then {m2} "
          f"particles", flush=True)
    runs = {}
    for mode in ("alloc", "noalloc"):
        with _ctx(golden_mesh(Z)) as ctx:
            _set_velocity(ctx, Z)
            step = (ctx.particles_step_host if mode == "alloc"
                    else ctx.particles_step_host_noalloc)
            p1 = _particles(rec)
            n1 = step(p1, t1, dt)
            p2 = _head(p1, m2)
            n2 = step(p2, t2, dt)
            runs[mode] = (p1, n1, p2, n2)
    a, b = runs["alloc"], runs["noalloc"]
    assert a[1] == b[1] and a[3] == b[3]
    assert _same(a[0], b[0]), "first call differs"
    assert _same(a[2], b[2]), "second (smaller) call differs"
    # the step moved the particles (the comparison is not of unchanged input)
    assert not np.array_equal(a[0]["x"], rec[:, 1])


def test_noalloc_particle_step_refusals():
    Z = golden("channel_n6")
    h, rec = Z["particle_hdr"], Z["in_particles"]
    with _ctx(golden_mesh(Z)) as ctx:
        p = _particles(rec[:10])
        with pytest.raises(S.SsxError) as e:    # tables not set up yet
            ctx.particles_step_host_noalloc(p, float(h[0]), float(h[1]))
        assert e.value.code == 1
        _set_velocity(ctx, Z)
        big = _particles(np.repeat(rec[:1], S.NOALLOC_MAX_PARTICLES + 1, 0))
        with pytest.raises(S.SsxError) as e:    # beyond the stated maximum
            ctx.particles_step_host_noalloc(big, float(h[0]), float(h[1]))
        assert e.value.code == 1
        assert np.array_equal(big["x"], np.repeat(rec[:1, 1],
                                                  len(big["x"])))
