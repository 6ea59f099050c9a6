"""FCG + p-multigrid on mid-size deformed cubes against the oracle (the C
restatement of solver.c / precon.c, run here with OpenMP on the host):
16^3 elements at n = 8 (2.1M local nodes) and 10^3 at n = 12 (1.7M),
Dirichlet walls, x' = x + 0.1 sin(pi x) sin(pi y) sin(pi z) on each
coordinate, the 16-mode Fourier source in weak form, tolerance 1e-10
relative in the multiplicity-weighted norm.

Bars: FP64 smoother -- iterations within one of the oracle's, solution
within 1e-8 relative; FP32 and BF16-factor smoothers (the FP64 outer FCG
reaches the same tolerance) -- iterations within one of the oracle's FP64
count, solution within 1e-7 relative of the oracle's."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from ssenss_amd.mesh import box_mesh  # noqa: E402
from conftest import rel  # noqa: E402
from oracle import pyoracle as P  # noqa: E402

pytestmark = pytest.mark.gpu

_CACHE = {}


def _source(x, y, z, seed=2024):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 9, size=(16, 3))
    a = rng.standard_normal(16)
    f = np.zeros_like(x)
    for m in range(16):
        f += a[m] * (np.sin(np.pi * k[m, 0] * x) * np.sin(np.pi * k[m, 1] * y)
                     * np.sin(np.pi * k[m, 2] * z))
    return f


def _oracle_case(n, m):
    """mesh, right-hand side and the oracle's FP64 FCG + p-MG solve"""
    key = (n, m)
    if key in _CACHE:
        return _CACHE[key]
    mesh = box_mesh(n, (m, m, m), deform=1, amp=0.1)
    P.set_threads(min(16, os.cpu_count() or 1))
    o = P.Oracle(mesh)
    _, rho, _, _ = P.gll_basis(n)
    w3 = (rho[None, None, :] * rho[None, :, None] * rho[:, None, None]).ravel()
    f = _source(mesh.x.ravel(), mesh.y.ravel(), mesh.z.ravel())
    b = o.mask(0, P.X_VELOCITY, o.dss(0, np.tile(w3, mesh.nel)
                                      * o.data(0, P.DETJ) * f))
    eps = 1e-10 * np.sqrt(o.dot(0, b, b))
    o.set_diag_choice(0)
    u, it, err, _ = o.fcg(P.OP_POISSON, P.X_VELOCITY, P.PC_MG, b,
                          np.zeros(mesh.ntot), eps, 500)
    o.close()
    _CACHE[key] = (mesh, b, eps, u, it)
    return _CACHE[key]


@pytest.mark.parametrize("n,m", [(8, 16), (12, 10)])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_fcg_mg_midsize_against_oracle(n, m, prec):
    mesh, b, eps, u_o, it_o = _oracle_case(n, m)
    print(f"[gpu] midsize {m}^3 n={n} prec {prec}: oracle {it_o} its",
          flush=True)
    ctx = S.Context(mesh)
    ctx.set_stream(torch.cuda.current_stream())
    with ctx:
        ctx.set_diag_choice(0)
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        f = torch.from_numpy(b).cuda()
        u = torch.zeros_like(f)
        st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, u, eps, 500)
        torch.cuda.synchronize()
        err = rel(u.cpu().numpy(), u_o)
    print(f"[gpu]   {st.iterations} its (oracle {it_o}), |u-u_o|/|u_o| "
          f"{err:.2e}, coarse {st.coarse_iters}")
    assert st.converged
    assert abs(st.iterations - it_o) <= 1
    assert err <= (1e-8 if prec == 0 else 1e-7)
