"""The sparse direct coarsest-level solve (SSX_OPT_COARSE_SPARSE): the
assembled coarse operator factored by the multifrontal Cholesky of
ssx_sdirect.c, solved on the device, in place of the coarse CG of the
reduced-precision V-cycle.  Against the dense direct solve (the same
system, inverted by rocSOLVER) and the coarse CG: all reach the FP64
tolerance, the two direct solves give the same outer iteration count and
solutions agree to the tolerance; two runs give identical bits.  Golden
meshes (pressure, whose pinned node replaces the null space, and Dirichlet
velocity, periodic included) and a 24^3 deformed cube whose coarsest level
(15,625 nodes) has separator levels split over many workgroups."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from ssenss_amd.mesh import box_mesh  # noqa: E402
from conftest import golden, golden_mesh, rel  # noqa: E402

pytestmark = pytest.mark.gpu


def _solve(mesh, f_host, var, mode, eps, prec=S.PREC_FP32):
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        ctx.set_diag_choice(-1 if var == S.PRESSURE else 0)
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        ctx.set_option(S.OPT_COARSE_DIRECT, 16384 if mode == "dense" else 0)
        ctx.set_option(S.OPT_COARSE_SPARSE, 1 << 20 if mode == "sparse" else 0)
        f = torch.from_numpy(np.ascontiguousarray(f_host)).cuda()
        out = []
        for _ in range(2 if mode == "sparse" else 1):
            u = torch.zeros_like(f)
            st = ctx.fcg(S.OP_POISSON, var, S.PC_MG, f.clone(), u, eps, 500)
            torch.cuda.synchronize()
            out.append((u.cpu().numpy(), st))
        return out


def _compare(name, mesh, f, var, eps):
    (u_cg, st_cg), = _solve(mesh, f, var, "cg", eps)
    (u_d, st_d), = _solve(mesh, f, var, "dense", eps)
    (u_s, st_s), (u_s2, _) = _solve(mesh, f, var, "sparse", eps)
    print(f"[gpu] {name} var {var}: CG {st_cg.iterations} its, dense "
          f"{st_d.iterations}, sparse {st_s.iterations}; |u_s - u_d| "
          f"{rel(u_s, u_d):.1e}, |u_s - u_cg| {rel(u_s, u_cg):.1e}",
          flush=True)
    assert st_s.converged and st_s.residual <= eps
    assert abs(st_s.iterations - st_d.iterations) <= 1
    assert st_s.iterations <= st_cg.iterations + 1
    assert st_s.coarse_iters == st_s.vcycles       # one solve per cycle
    assert rel(u_s, u_d) <= 1e-7 and rel(u_s, u_cg) <= 1e-7
    assert np.array_equal(u_s, u_s2)


@pytest.mark.parametrize("name", ["channel_n6", "cube2_n8", "periodic3_n6"])
@pytest.mark.parametrize("var", [S.PRESSURE, S.X_VELOCITY])
def test_sparse_coarse_golden(name, var):
    Z = golden(name)
    _compare(name, golden_mesh(Z), Z["in_f"], var, Z["ref_scalars"][0])


def test_sparse_coarse_split_levels():
    mesh = box_mesh(4, (24, 24, 24), deform=1, amp=0.1)
    rng = np.random.default_rng(24)
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        assert ctx.level_n(ctx.nlevels - 1) == 2
        m = torch.from_numpy(ctx.level_data(0, S.DATA_MASS)).cuda()
        b = torch.from_numpy(rng.standard_normal(mesh.ntot)).cuda() * m
        ctx.dss(0, b)
        ctx.mask(0, S.X_VELOCITY, b)
        eps = 1e-10 * np.sqrt(ctx.dot(0, b, b))
        f = b.cpu().numpy()
    _compare("cube24_n4", mesh, f, S.X_VELOCITY, eps)


def test_sparse_coarse_leaves_fp64_cycle_alone():
    Z = golden("channel_n6")
    mesh = golden_mesh(Z)
    eps = Z["ref_scalars"][0]
    (u0, st0), = _solve(mesh, Z["in_f"], S.PRESSURE, "cg", eps, S.PREC_FP64)
    (u1, st1), _ = _solve(mesh, Z["in_f"], S.PRESSURE, "sparse", eps,
                          S.PREC_FP64)
    assert st0.iterations == st1.iterations
    assert np.array_equal(u0, u1)
