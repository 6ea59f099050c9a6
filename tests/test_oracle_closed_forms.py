"""Checks of the oracle against what the mathematics fixes, independent of
the reference: closed-form GLL nodes, quadrature and differentiation
identities, polynomial reproduction of the transfer matrices, the affine
geometry closed form, a brute-force quadrature assembly of the element
stiffness, symmetry of the assembled operator, and textbook PCG on the
explicitly assembled unique-node system.  Each targets a plausible mistake
(dropped term, wrong index, transposed operand) that golden vectors on a few
meshes could miss."""
import numpy as np
import pytest

from oracle import pyoracle as P
from ssenss_amd.mesh import Mesh

FIVE = np.sqrt(3.0 / 7.0)


def test_gll_closed_forms():
    known = {2: [-1, 1], 3: [-1, 0, 1], 4: [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1],
             5: [-1, -FIVE, 0, FIVE, 1]}
    for n, pts in known.items():
        xi, rho, _, _ = P.gll_basis(n)
        assert np.allclose(xi, pts, atol=1e-15, rtol=0)
    for n in range(2, 17):
        xi, rho, d, _ = P.gll_basis(n)
        D = d.T                                 # D[i, j] = l_j'(xi_i)
        assert abs(rho.sum() - 2.0) <= 1e-14
        assert np.abs(D @ np.ones(n)).max() <= 1e-11 * n * n
        assert np.abs(D @ xi - 1.0).max() <= 1e-11 * n * n
        # summation by parts: rho_i D_ij + rho_j D_ji = e_N e_N - e_0 e_0
        S = rho[:, None] * D + (rho[:, None] * D).T
        E = np.zeros((n, n))
        E[0, 0], E[-1, -1] = -1, 1
        assert np.abs(S - E).max() <= 1e-12 * n
        # quadrature exact to degree 2n-3
        for k in range(2 * n - 2):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(rho @ xi ** k - exact) <= 1e-13


def test_interp_reproduces_polynomials():
    for nf in range(3, 17):
        nc = max(2, (nf - 1) // 2 + 1)
        J = P.interp_matrix(nc, nf)             # J[p, i] = l_p(xi_f_i)
        xc, _, _, _ = P.gll_basis(nc)
        xf, _, _, _ = P.gll_basis(nf)
        assert np.abs(J.sum(axis=0) - 1).max() <= 1e-13
        for k in range(nc):
            assert np.abs(J.T @ xc ** k - xf ** k).max() <= 1e-12
        # equal orders give the identity exactly
        assert np.array_equal(P.interp_matrix(nf, nf), np.eye(nf))


def single_element(n, x, y, z):
    nb = np.full((1, 6), -1, np.int32)
    return Mesh(n, x[None], y[None], z[None], nb,
                np.full((1, 12), -1, np.int32), np.zeros((1, 12), np.int32),
                np.full((1, 8), -1, np.int32), np.zeros((1, 8), np.int32))


def gll_numpy(n):
    """GLL nodes/weights/derivative matrix from numpy's Legendre module."""
    N = n - 1
    c = np.zeros(N + 1)
    c[N] = 1
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    x = np.concatenate([[-1.0], np.sort(inner), [1.0]])
    PN = np.polynomial.legendre.legval(x, c)
    w = 2.0 / (N * (N + 1) * PN ** 2)
    # barycentric derivative matrix
    lam = np.array([1.0 / np.prod([x[j] - x[m] for m in range(n) if m != j])
                    for j in range(n)])
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = lam[j] / lam[i] / (x[i] - x[j])
        D[i, i] = -D[i].sum()
    return x, w, D


def test_affine_geometry_closed_form():
    n, h = 5, 0.3
    xi, _, _, _ = P.gll_basis(n)
    t = 0.5 * (1 + xi) * h
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    o = P.Oracle(single_element(n, x + 1.0, y - 2.0, z + 0.5))
    assert np.allclose(o.data(0, P.DETJ), (h / 2) ** 3, rtol=1e-13)
    for a, v in zip(range(6), [1, 0, 0, 1, 0, 1]):
        assert np.allclose(o.data(0, a), v * (2 / h) ** 2, rtol=1e-12,
                           atol=1e-12 * (2 / h) ** 2)


@pytest.mark.parametrize("n", [3, 4, 6])
def test_element_stiffness_brute_force(n):
    """A_e = sum_q w_q detJ_q grad(phi_a).grad(phi_b) assembled from
    numpy's own basis on a curved element; compared with the oracle's Ax."""
    xg, w, D = gll_numpy(n)
    rng = np.random.default_rng(n)
    Z, Y, X = np.meshgrid(xg, xg, xg, indexing="ij")
    s = 0.08
    x = X + s * np.sin(1.3 * Y + 0.4) * np.cos(Z)
    y = Y + s * np.sin(0.7 * X + Z)
    z = Z + s * np.cos(X - 0.5 * Y)
    m = n ** 3
    I = np.eye(n)
    Dx = np.kron(I, np.kron(I, D))        # d/dxi   on [k][j][i]
    Dy = np.kron(I, np.kron(D, I))        # d/deta
    Dz = np.kron(D, np.kron(I, I))        # d/dzeta
    W = np.kron(w, np.kron(w, w))
    c = [x.ravel(), y.ravel(), z.ravel()]
    J = np.zeros((m, 3, 3))
    for a in range(3):
        for b, Db in enumerate((Dx, Dy, Dz)):
            J[:, a, b] = Db @ c[a]
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)                  # Ji[q, b, a] = d xi_b / d x_a
    K = np.zeros((m, m))
    Dref = [Dx, Dy, Dz]
    for a in range(3):                     # physical gradient component
        Ga = sum(Ji[:, b, a][:, None] * Dref[b] for b in range(3))
        K += Ga.T @ ((W * det)[:, None] * Ga)
    o = P.Oracle(single_element(n, x, y, z))
    u = rng.standard_normal(m)
    assert np.linalg.norm(o.ax(0, P.OP_POISSON, u) - K @ u) \
        <= 1e-12 * np.linalg.norm(K @ u)
    assert np.abs(o.ax(0, P.OP_POISSON, np.ones(m))).max() <= 1e-12 * np.abs(K).max()


def box(n, m, bc=-1, amp=0.1):
    from ssenss_amd.mesh import box_mesh
    return box_mesh(n, m, bc=(bc,) * 6, deform=1 if amp else 0, amp=amp,
                    dt=11 / 60, visc=1.0)


def unique_map(o):
    """local -> unique node numbering from the DSS of the one-hot sums"""
    nt = o.ntot()
    ids = np.arange(nt, dtype=np.float64)
    # the smallest local index of each node: DSS of 2^-id style is not exact;
    # use coordinates (non-periodic meshes: unique positions)
    xyz = np.stack([o.data(0, P.X), o.data(0, P.Y), o.data(0, P.Z)], 1)
    key = np.round(xyz * 1e9).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    return inv.ravel(), ids


def test_assembled_operator_symmetry():
    o = P.Oracle(box(5, (2, 2, 2)))
    rng = np.random.default_rng(3)
    x = o.dss_average(0, rng.standard_normal(o.ntot()))
    y = o.dss_average(0, rng.standard_normal(o.ntot()))
    for op in (P.OP_POISSON, P.OP_HELMHOLTZ):
        a = o.dot(0, x, o.opeval(0, op, P.PRESSURE, y))
        b = o.dot(0, y, o.opeval(0, op, P.PRESSURE, x))
        assert abs(a - b) <= 1e-12 * max(abs(a), 1)


def test_cg_equals_textbook_pcg_on_assembled_system():
    """solver.c:33-81 on E-vectors == textbook Jacobi-PCG on the unique-node
    system A_u = Q^T A_L Q restricted to interior nodes."""
    o = P.Oracle(box(3, (2, 2, 2)))
    nt = o.ntot()
    uid, _ = unique_map(o)
    nu = uid.max() + 1
    Q = np.zeros((nt, nu))
    Q[np.arange(nt), uid] = 1
    AL = np.column_stack([o.ax(0, P.OP_HELMHOLTZ, e) for e in np.eye(nt)])
    Au = Q.T @ AL @ Q
    mask = o.data(0, P.MASK)
    interior = np.zeros(nu, bool)
    interior[uid[mask == 1]] = True
    dinv = np.zeros(nu)
    dinv[uid] = o.data(0, P.DIAG_H)
    rng = np.random.default_rng(5)
    b = o.dss(0, rng.standard_normal(nt))
    bu = np.zeros(nu)
    bu[uid] = b                                 # continuous after DSS
    A = Au[np.ix_(interior, interior)]
    Mi = dinv[interior]
    rhs = bu[interior]
    eps = 1e-10 * np.linalg.norm(rhs)
    # textbook PCG with the reference's loop structure
    xk = np.zeros_like(rhs)
    r = rhs.copy()
    z = Mi * r
    p = z.copy()
    rz = r @ z
    it = 0
    while True:
        q = A @ p
        al = rz / (p @ q)
        xk += al * p
        r -= al * q
        z = Mi * r
        rz, rz0 = r @ z, rz
        p = z + (rz / rz0) * p
        it += 1
        if np.sqrt(r @ r) <= eps or it >= 500:
            break
    sol, it_o, err, _ = o.cg(0, P.OP_HELMHOLTZ, P.X_VELOCITY, b, np.zeros(nt),
                             eps, 500)
    assert abs(it_o - it) <= 1
    su = np.zeros(nu)
    su[uid] = sol
    assert np.linalg.norm(su[interior] - xk) <= 1e-8 * np.linalg.norm(xk)


def test_fcg_with_jacobi_equals_pcg_iterations():
    o = P.Oracle(box(4, (2, 2, 2)))
    rng = np.random.default_rng(9)
    b = o.dss(0, rng.standard_normal(o.ntot()))
    eps = 1e-10 * np.sqrt(o.dot(0, b, b))
    _, it_cg, _, _ = o.cg(0, P.OP_HELMHOLTZ, P.X_VELOCITY, b,
                          np.zeros(o.ntot()), eps, 1000)
    _, it_f, _, _ = o.fcg(P.OP_HELMHOLTZ, P.X_VELOCITY, P.PC_JACOBI, b,
                          np.zeros(o.ntot()), eps, 1000)
    assert abs(it_cg - it_f) <= 1


def test_unique_node_count():
    for m, n in (((2, 3, 2), 4), ((3, 3, 3), 5)):
        o = P.Oracle(box(n, m))
        one = np.ones(o.ntot())
        exp = np.prod([mi * (n - 1) + 1 for mi in m])
        assert o.dot(0, one, one) == exp
    o = P.Oracle(box(5, (3, 3, 3), bc=0, amp=0))
    assert o.dot(0, np.ones(o.ntot()), np.ones(o.ntot())) == (3 * 4) ** 3
