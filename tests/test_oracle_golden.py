"""The oracle (CPU restatement) against golden vectors written by the
compiled reference (tests/golden/make_golden.py).  Tolerances: bit-exact for
integer-valued DSS and the multiplicity/mask arrays; 1e-14 relative for
pointwise set-up quantities; 1e-12 relative for operator outputs (BLAS
summation order is not fixed by the reference); iteration counts within one
(they match exactly on these cases)."""
import numpy as np
import pytest

from conftest import rel
from oracle import pyoracle as P


def test_basis_matches_reference():
    B = np.load("tests/golden/basis.npz")
    for n in range(2, 17):
        xi, rho, d, ds = P.gll_basis(n)
        assert np.abs(xi - B[f"xi_{n}"]).max() <= 1e-15, n
        assert np.abs(rho - B[f"rho_{n}"]).max() <= 1e-15, n
        # D scales like n^2; entries agree to a few ulps of their size
        assert np.abs(d.ravel() - B[f"d_{n}"]).max() <= 2e-14 * n * n, n
        assert np.abs(ds.ravel() - B[f"dsvv_{n}"]).max() <= 2e-14 * n * n, n


def test_interp_matrices_match_reference():
    B = np.load("tests/golden/basis.npz")
    for nf in range(3, 17):
        nc = max(2, (nf - 1) // 2 + 1)
        assert np.abs(P.interp_matrix(nc, nf).ravel() - B[f"jcf_{nf}"]).max() \
            <= 1e-14
        assert np.abs(P.interp_matrix(nf, nc).ravel() - B[f"jfc_{nf}"]).max() \
            <= 1e-13


def test_geometry_and_level_data(case):
    name, Z, mesh = case
    o = P.Oracle(mesh)
    nt = o.ntot()
    g = Z["ref_geom"].reshape(7, nt)
    for a in range(6):
        assert rel(o.data(0, a), g[a]) <= 1e-14, (name, a)
    assert rel(o.data(0, P.DETJ), g[6]) <= 1e-14
    assert list(Z["levels"]) == [o.level_n(l) for l in range(o.nlevels)]
    for l in range(o.nlevels):
        L = Z[f"ref_level{l}"].reshape(4, -1)
        assert rel(o.data(l, P.DIAG_P), L[0]) <= 1e-13, (name, l)
        assert rel(o.data(l, P.DIAG_H), L[1]) <= 1e-13, (name, l)
        assert np.array_equal(o.data(l, P.SCALE), L[2]), (name, l)
        assert np.array_equal(o.data(l, P.MASK), L[3]), (name, l)
    assert rel(o.data(0, P.INVMASS), Z["ref_invmass"]) <= 1e-14


def test_operators(case):
    name, Z, mesh = case
    o = P.Oracle(mesh)
    u = Z["in_u"]
    assert rel(o.ax(0, P.OP_POISSON, u), Z["ref_ax_poisson"]) <= 1e-12
    assert rel(o.ax(0, P.OP_HELMHOLTZ, u), Z["ref_ax_helmholtz"]) <= 1e-12
    assert rel(o.opeval(0, P.OP_POISSON, P.PRESSURE, u),
               Z["ref_opeval_pressure"]) <= 1e-12
    assert rel(o.opeval(0, P.OP_HELMHOLTZ, P.X_VELOCITY, u),
               Z["ref_opeval_helm_vel"]) <= 1e-12


def test_dss_bit_exact(case):
    name, Z, mesh = case
    o = P.Oracle(mesh)
    nt = o.ntot()
    q = np.arange(nt, dtype=np.uint64)
    iv = ((q * np.uint64(2654435761)) % np.uint64(1000003)).astype(np.float64)
    assert np.array_equal(o.dss(0, iv), Z["ref_dss_int"])
    # same add order as the reference: random fields agree bit for bit too
    assert np.array_equal(o.dss(0, Z["in_u"]), Z["ref_dss_u"])
    assert np.array_equal(o.dss_average(0, Z["in_u"]), Z["ref_dss_avg"])


def test_dots_nullspace_transfers(case):
    name, Z, mesh = case
    o = P.Oracle(mesh)
    u = Z["in_u"]
    one = np.ones(o.ntot())
    d = Z["ref_dots"]
    assert abs(o.dot(0, u, u) / d[0] - 1) <= 1e-13
    assert abs(o.dot(0, u, one) - d[1]) <= 1e-12 * np.sqrt(d[0] * d[2])
    assert o.dot(0, one, one) == d[2]
    assert rel(o.nullspace(0, P.PRESSURE, u), Z["ref_nullspace"]) <= 1e-14
    if o.nlevels > 1:
        ref = o.dss_average(1, Z["ref_restrict_raw"])
        assert rel(o.restrict(1, u), ref) <= 1e-13
        assert rel(o.prolong(0, Z["in_v1"]), Z["ref_prolong"]) <= 1e-13


def test_solvers_iterations_and_solutions(case):
    name, Z, mesh = case
    o = P.Oracle(mesh)
    nt = o.ntot()
    b = Z["in_f"]
    sc = Z["ref_scalars"]
    eps = sc[0]
    sol, it, err, _ = o.cg(0, P.OP_HELMHOLTZ, P.X_VELOCITY, b, np.zeros(nt),
                           eps, 5000)
    assert abs(it - int(sc[1])) <= 1 and err <= eps
    assert rel(sol, Z["ref_cg_helm"]) <= 1e-9
    sol, it, err, _ = o.fcg(P.OP_POISSON, P.PRESSURE, P.PC_MG, b,
                            np.zeros(nt), eps, 5000)
    assert abs(it - int(sc[3])) <= 1 and err <= eps
    assert rel(sol, Z["ref_fcg_pressure"]) <= 1e-9
    assert rel(o.mg_precon(P.OP_POISSON, P.PRESSURE, eps, b),
               Z["ref_mg_pressure"]) <= 1e-10
    o.set_diag_choice(0)
    sol, it, err, _ = o.fcg(P.OP_POISSON, P.X_VELOCITY, P.PC_MG, b,
                            np.zeros(nt), eps, 5000)
    assert abs(it - int(sc[5])) <= 1 and err <= eps
    assert rel(sol, Z["ref_fcg_dirichlet"]) <= 1e-9
    assert rel(o.mg_precon(P.OP_POISSON, P.X_VELOCITY, eps,
                           o.mask(0, P.X_VELOCITY, b)),
               Z["ref_mg_dirichlet"]) <= 1e-9


@pytest.mark.parametrize("threads", [1, 4])
def test_oracle_thread_count_does_not_change_results(threads):
    from conftest import golden, golden_mesh
    Z = golden("cube2_n8")
    P.set_threads(threads)
    try:
        o = P.Oracle(golden_mesh(Z))
        assert np.array_equal(o.dss(0, Z["in_u"]), Z["ref_dss_u"])
        a = o.ax(0, P.OP_POISSON, Z["in_u"])
    finally:
        P.set_threads(1)
    o2 = P.Oracle(golden_mesh(Z))
    assert np.array_equal(a, o2.ax(0, P.OP_POISSON, Z["in_u"]))


@pytest.mark.parametrize("name", ["box2_n5", "periodic3_n6", "channel_n6"])
def test_particles_match_reference(name):
    """integrate_particle_equations with injected particles: owning
    elements bit-exact; states bit-exact where the reference interpolates
    with plain loops (n > 5), to rounding where it calls BLAS (n <= 5)."""
    from conftest import golden, golden_mesh
    Z = golden(name)
    o = P.Oracle(golden_mesh(Z))
    h = Z["particle_hdr"]
    out, re, passes = o.particles(Z["in_vel"], h[2:], Z["in_particles"],
                                  h[0], h[1])
    ref = Z["ref_particles"].reshape(-1, 8)
    assert re.sum() == 0 and passes >= 2
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert (out[:, 0] != Z["in_particles"][:, 0]).sum() > 20   # hops happen
    if int(Z["n"]) > 5:
        assert np.array_equal(out, ref)
    else:
        assert np.abs(out[:, 1:] - ref[:, 1:]).max() <= 1e-13


def test_point_interpolation_closed_form():
    """full Lagrange (n <= 5) and the 5-point stencil (n > 5) reproduce
    polynomials of degree <= min(n-1, 4) exactly"""
    rng = np.random.default_rng(4)
    for n in (3, 5, 6, 8, 12):
        xi, _, _, _ = P.gll_basis(n)
        deg = min(n - 1, 4)
        c = rng.standard_normal((deg + 1, deg + 1, deg + 1))
        Z, Y, X = np.meshgrid(xi, xi, xi, indexing="ij")
        f = np.polynomial.polynomial.polyval3d(X, Y, Z, c)
        for _ in range(20):
            x, y, z = rng.uniform(-1, 1, 3)
            ref = np.polynomial.polynomial.polyval3d(x, y, z, c)
            got = P.interp_point(n, xi, f.ravel(), x, y, z)
            assert abs(got - ref) <= 1e-12 * max(1, abs(ref)), (n, got, ref)
