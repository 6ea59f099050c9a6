"""The direct coarsest-level solve (SSX_OPT_COARSE_DIRECT): the assembled
coarse operator inverted once by Cholesky, used in place of the coarse CG of
the reduced-precision V-cycle.  Against the same solve with the coarse CG:
both reach the FP64 tolerance, the solutions agree to the level the
tolerance fixes, and the direct solve takes no more outer iterations (it
solves the coarse level exactly).  With the FP64 smoother the option
changes nothing (the reference's coarse CG runs), bit for bit.  Pressure
(singular: the constant null space) and Dirichlet velocity cases."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from conftest import golden, golden_mesh, rel  # noqa: E402

pytestmark = pytest.mark.gpu


def _solve(mesh, f_host, var, prec, direct, eps):
    with S.Context(mesh) as ctx:
        ctx.set_stream(torch.cuda.current_stream())
        ctx.set_diag_choice(-1 if var == S.PRESSURE else 0)
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        ctx.set_option(S.OPT_COARSE_DIRECT, 16384 if direct else 0)
        ctx.set_option(S.OPT_COARSE_SPARSE, 0)       # the CG otherwise
        f = torch.from_numpy(np.ascontiguousarray(f_host)).cuda()
        u = torch.zeros_like(f)
        st = ctx.fcg(S.OP_POISSON, var, S.PC_MG, f, u, eps, 500)
        torch.cuda.synchronize()
        return u.cpu().numpy(), st


@pytest.mark.parametrize("name", ["channel_n6", "cube2_n8", "periodic3_n6"])
@pytest.mark.parametrize("var", [S.PRESSURE, S.X_VELOCITY])
def test_direct_coarse_solve(name, var):
    Z = golden(name)
    mesh = golden_mesh(Z)
    eps = Z["ref_scalars"][0]
    f = Z["in_f"]
    u_cg, st_cg = _solve(mesh, f, var, S.PREC_FP32, False, eps)
    u_d, st_d = _solve(mesh, f, var, S.PREC_FP32, True, eps)
    err = rel(u_d, u_cg)
    print(f"[gpu] {name} var {var}: coarse CG {st_cg.iterations} its "
          f"({st_cg.coarse_iters} coarse), direct {st_d.iterations} its "
          f"({st_d.coarse_iters} coarse), |u_d - u_cg|/|u_cg| {err:.2e}")
    assert st_cg.converged and st_d.converged
    assert st_d.residual <= eps
    assert st_d.iterations <= st_cg.iterations + 1
    assert st_d.coarse_iters == st_d.vcycles       # one solve per cycle
    assert err <= 1e-7


def test_direct_coarse_leaves_fp64_cycle_alone():
    Z = golden("channel_n6")
    mesh = golden_mesh(Z)
    eps = Z["ref_scalars"][0]
    u0, st0 = _solve(mesh, Z["in_f"], S.PRESSURE, S.PREC_FP64, False, eps)
    u1, st1 = _solve(mesh, Z["in_f"], S.PRESSURE, S.PREC_FP64, True, eps)
    assert st0.iterations == st1.iterations
    assert np.array_equal(u0, u1)
