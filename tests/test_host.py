"""Host-side logic of the product library (no GPU): the C ABI exports every
declared symbol, the mesh generator follows the reference's grid.h5
conventions, and the host assembly (shared-node groups, geometry) matches
the compiled reference bit for bit / to rounding."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ssenss_amd as S
from ssenss_amd.mesh import box_mesh, NOSLIP, PERIODIC
from conftest import ROOT, rel


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ssx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ssx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = S.load()
    declared = header_functions()
    assert declared, "no declarations parsed"
    assert sorted(S.EXPORTED_SYMBOLS) == declared
    raw = C.CDLL(S.library_path())
    for name in declared:
        assert hasattr(raw, name), name
    assert L.ssx_version().decode().startswith("ssx")


def test_generator_matches_reference_example_grid():
    """examples/grid.h5 of the reference: 2x2x2 no-slip box on [-1,1]^3,
    written at n = 2 (misc/rectangular_grid.py)."""
    E = np.load(os.path.join(ROOT, "tests", "golden", "example_grid.npz"))
    m = box_mesh(2, (2, 2, 2), lo=(-1, -1, -1), hi=(1, 1, 1),
                 bc=(NOSLIP,) * 6)
    assert np.array_equal(m.neigh, E["neigh"])
    assert np.array_equal(m.edge_neigh, E["edge_neigh"])
    assert np.array_equal(m.edge_size, E["edge_size"])
    assert np.array_equal(m.corn_neigh, E["corn_neigh"])
    assert np.array_equal(m.corn_size, E["corn_size"])
    assert np.array_equal(E["rot"], np.zeros_like(E["rot"]))
    assert np.array_equal(m.x.reshape(8, 8), E["x"])
    assert np.array_equal(m.y.reshape(8, 8), E["y"])
    assert np.array_equal(m.z.reshape(8, 8), E["z"])


def host_dss(mesh, n, u):
    u = np.array(u, np.float64, copy=True)
    ms = mesh.struct()
    S.native.check(S.load().ssx_dss_host(
        C.byref(ms), n, u.ctypes.data_as(C.POINTER(C.c_double))))
    return u


def test_host_dss_bit_exact_against_reference(case):
    name, Z, mesh = case
    nt = mesh.ntot
    q = np.arange(nt, dtype=np.uint64)
    iv = ((q * np.uint64(2654435761)) % np.uint64(1000003)).astype(np.float64)
    assert np.array_equal(host_dss(mesh, mesh.n, iv), Z["ref_dss_int"])
    assert np.array_equal(host_dss(mesh, mesh.n, Z["in_u"]), Z["ref_dss_u"])
    # coarse levels: multiplicity 1/DSS(1) equals the reference's scaling
    for l, nl in enumerate(Z["levels"]):
        ref = Z[f"ref_level{l}"].reshape(4, -1)[2]
        one = np.ones(mesh.nel * int(nl) ** 3)
        assert np.array_equal(1.0 / host_dss(mesh, int(nl), one), ref)


def gather_dss(mesh, n, u):
    u = np.array(u, np.float64, copy=True)
    ms = mesh.struct()
    S.native.check(S.load().ssx_dss_gather_host(
        C.byref(ms), n, u.ctypes.data_as(C.POINTER(C.c_double))))
    return u


def test_gather_dss_bit_exact_against_reference(case):
    """The gather form (per-copy partner lists, used by the fused CG update)
    against the reference's own DSS output, bit for bit."""
    name, Z, mesh = case
    assert np.array_equal(gather_dss(mesh, mesh.n, Z["in_u"]), Z["ref_dss_u"])


@pytest.mark.parametrize("n,dims,bc", [
    (8, (5, 4, 3), (0, 0, 0, 0, 0, 0)),
    (8, (3, 4, 5), (-1, -1, -1, -1, -1, -1)),
    (4, (6, 5, 4), (0, 0, -1, -1, 0, 0)),
    (2, (6, 5, 4), (-1, -1, 0, 0, -2, -2)),
    (5, (3, 3, 3), (0, 0, 0, 0, 0, 0)),
    (7, (1, 2, 3), (-1, -1, -1, -1, -1, -1)),
])
def test_gather_dss_equals_group_dss(n, dims, bc):
    """Every level order and boundary mix: gather form == group form, bit
    for bit, on values whose sums round differently in every add order."""
    mesh = S.box_mesh(n, dims, bc=bc, deform=1, amp=0.1)
    rng = np.random.default_rng(n * 100 + dims[0])
    u = rng.standard_normal(mesh.ntot) * np.exp(rng.uniform(-20, 20,
                                                            mesh.ntot))
    g = gather_dss(mesh, n, u)
    assert np.array_equal(g, host_dss(mesh, n, u))
    assert not np.array_equal(g, u)


def test_host_geometry_against_reference(case):
    name, Z, mesh = case
    n, nt = mesh.n, mesh.ntot
    gf = np.zeros(6 * nt)
    mass = np.zeros(nt)
    ms = mesh.struct()
    dp = C.POINTER(C.c_double)
    S.native.check(S.load().ssx_geometry_host(
        C.byref(ms), gf.ctypes.data_as(dp), mass.ctypes.data_as(dp)))
    B = np.load(os.path.join(ROOT, "tests", "golden", "basis.npz"))
    rho = B[f"rho_{n}"]
    w3 = np.tile((rho[None, None, :] * rho[None, :, None]
                  * rho[:, None, None]).ravel(), mesh.nel)
    g = Z["ref_geom"].reshape(7, nt)
    assert rel(mass, w3 * g[6]) <= 1e-13
    gfe = gf.reshape(mesh.nel, 6, -1)
    ref = np.stack([(g[a] * w3 * g[6]).reshape(mesh.nel, -1)
                    for a in range(6)], axis=1)
    # all six factors together: off-diagonal ones vanish on affine meshes
    assert rel(gfe, ref) <= 1e-13


def test_unsupported_connectivity_is_rejected():
    # two elements per periodic direction: the reference's first-back-face
    # search wires the wrong faces (init.c:78-81); the engine refuses it
    m = box_mesh(3, (2, 3, 3), bc=(PERIODIC,) * 6)
    with pytest.raises(S.SsxError) as e:
        host_dss(m, 3, np.ones(m.ntot))
    assert e.value.code == 5


def test_bad_arguments():
    with pytest.raises(S.SsxError) as e:
        box_mesh(1, (2, 2, 2))
    assert e.value.code == 1
    with pytest.raises(S.SsxError):
        box_mesh(17, (1, 1, 1))


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    m = box_mesh(3, (2, 2, 2))
    with pytest.raises(S.SsxError) as e:
        S.Context(m)
    assert e.value.code == 6


def test_deformation_keeps_shared_nodes_identical():
    m = box_mesh(6, (3, 2, 2), deform=1, amp=0.1)
    for a in (m.x, m.y, m.z):
        # DSS-average of a continuous field is the field itself
        v = a.ravel()
        mult = host_dss(m, 6, np.ones(m.ntot))
        assert np.allclose(host_dss(m, 6, v) / mult, v, rtol=0, atol=1e-15)
