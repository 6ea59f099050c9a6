"""The Navier-Stokes time step (ssx_ns_*, navier_stokes.c:224-235) on the
GPU against the compiled reference's own integrate_navier_stokes
(fixtures tests/golden/ns_*.npz written by oracle/_ref/ref_harness ns):
u, v, w and p after each stored step.

Bars: velocity within 1e-11 relative and pressure within 1e-9 relative of
the reference's (both runs solve to 1e-10 in the weighted norm, so the
fields agree to the solve tolerance, not to rounding; measured 3e-14 and
3e-12 after 10 steps of the 4^3 n = 8 Taylor-Green case); the same with
the velocity solves batched or one by one."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from conftest import golden, golden_mesh, rel  # noqa: E402

pytestmark = pytest.mark.gpu

VEL_TOL, P_TOL = 1e-11, 1e-9


def _run(name, batch):
    Z = golden(name)
    mesh = golden_mesh(Z)
    nt = mesh.ntot
    ctx = S.Context(mesh)
    ctx.set_stream(torch.cuda.current_stream())
    g = tuple(float(x) for x in Z["g"])
    out = {}
    with ctx, S.NavierStokes(ctx, float(Z["dt"]), float(Z["visc"]),
                             float(Z["dens"]), 0.0, g, batch) as ns:
        vel = Z["in_vel"]
        levels = [[torch.from_numpy(np.ascontiguousarray(
            vel[(3 * l + c) * nt:(3 * l + c + 1) * nt])).cuda()
            for c in range(3)] for l in range(3)]
        ns.init(levels)
        for s in range(1, int(max(Z["steps"])) + 1):
            st = ns.step()
            if s in Z["steps"]:
                f = [ns.read(k).cpu().numpy() for k in range(4)]
                out[s] = (f, st)
    return Z, nt, out


@pytest.mark.parametrize("name", ["ns_tgv3_n6", "ns_channel_n6",
                                  "ns_tgv4_n8"])
@pytest.mark.parametrize("batch", [True, False])
def test_ns_step_against_reference(name, batch):
    print(f"[gpu] ns {name} batch={batch}", flush=True)
    Z, nt, out = _run(name, batch)
    for s, (f, st) in out.items():
        ref = Z[f"ref_step{s}"]
        eu = rel(np.concatenate(f[:3]), ref[:3 * nt])
        ep = rel(f[3], ref[3 * nt:])
        print(f"[gpu]   step {s}: |u-u_ref|/|u_ref| {eu:.2e}, |p-p_ref|/"
              f"|p_ref| {ep:.2e}; pressure FCG {st[0].iterations} its, "
              f"velocity CG {[x.iterations for x in st[1:]]}")
        assert all(x.converged for x in st)
        assert eu <= VEL_TOL
        assert ep <= P_TOL
