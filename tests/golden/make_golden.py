#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Runs ONLY in the build container, where the SSENSS sources are mounted:
  make -C oracle ref       # oracle/_ref/libssenss_ref.so + ref_harness
  python tests/golden/make_golden.py

For each mesh case the inputs (mesh arrays, a random field u, a right-hand
side b, a random coarse field v1) are generated here with committed seeds,
handed to oracle/_ref/ref_harness (the reference's own code), and the
reference's outputs are stored with the inputs in <case>.npz (float64, int32;
no pickles).  manifest.json lists sha256, sizes, seeds and iteration counts.

The right-hand side is the weak form the reference assembles for a nodal
source (navier_stokes.c:187-200 pattern): b = DSS(w * detJ * f), with
f = sum_m a_m sin(pi k1 x) sin(pi k2 y) sin(pi k3 z), 16 modes, k in 1..8,
a ~ N(0,1), then masked on Dirichlet faces by the solvers themselves.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from ssenss_amd.mesh import box_mesh, NOSLIP, PERIODIC  # noqa: E402
from oracle import pyoracle  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
EXAMPLE_GRID = "/root/reference/examples/grid.h5"
H5DUMP = "/opt/conda/bin/h5dump"

CASES = [
    # name, n, m, lo, hi, bc, deform, amp, dt, visc, seed
    dict(name="box2_n5", n=5, m=(2, 2, 2), lo=(-1, -1, -1), hi=(1, 1, 1),
         bc=(NOSLIP,) * 6, deform=1, amp=0.1, dt=11 / 60, visc=1.0, seed=11,
         particles=1000),
    dict(name="cube2_n8", n=8, m=(2, 2, 2), lo=(0, 0, 0), hi=(1, 1, 1),
         bc=(NOSLIP,) * 6, deform=1, amp=0.1, dt=11 / 600, visc=1.0, seed=12),
    dict(name="periodic3_n6", n=6, m=(3, 3, 3), lo=(0, 0, 0),
         hi=(2 * np.pi,) * 3, bc=(PERIODIC,) * 6, deform=0, amp=0.0,
         dt=1e-3, visc=1 / 1600, seed=13, particles=1000),
    dict(name="channel_n6", n=6, m=(3, 2, 3), lo=(0, -1, 0),
         hi=(2 * np.pi, 1, np.pi), bc=(PERIODIC, PERIODIC, NOSLIP, NOSLIP,
                                       PERIODIC, PERIODIC),
         deform=2, amp=2.0, dt=2e-3, visc=1 / 180, seed=14, particles=1000),
    dict(name="cube2_n12", n=12, m=(2, 2, 2), lo=(0, 0, 0), hi=(1, 1, 1),
         bc=(NOSLIP,) * 6, deform=1, amp=0.1, dt=11 / 600, visc=1.0, seed=15),
]


def smooth_source(x, y, z, seed, lo, hi):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 9, size=(16, 3))
    a = rng.standard_normal(16)
    L = np.asarray(hi, float) - np.asarray(lo, float)
    u = (x - lo[0]) / L[0]
    v = (y - lo[1]) / L[1]
    w = (z - lo[2]) / L[2]
    f = np.zeros_like(x)
    for m in range(16):
        f += a[m] * (np.sin(np.pi * k[m, 0] * u) * np.sin(np.pi * k[m, 1] * v)
                     * np.sin(np.pi * k[m, 2] * w))
    return f


def weak_rhs(mesh, f_nodal):
    """b = DSS(w * detJ * f) computed with the oracle's geometry."""
    orc = pyoracle.Oracle(mesh)
    n = mesh.n
    _, rho, _, _ = pyoracle.gll_basis(n)
    w3 = (rho[None, None, :] * rho[None, :, None] * rho[:, None, None]).ravel()
    detj = orc.data(0, pyoracle.DETJ)
    b = np.tile(w3, mesh.nel) * detj * f_nodal.ravel()
    return orc.dss(0, b)


PARTICLE_T, PARTICLE_DT, PARTICLE_SCALES = 1.0, 0.05, (0.8, 0.9, 1.0)


def particle_inputs(mesh, lo, hi, count, seed):
    """A smooth periodic velocity field (three history levels are this
    field times PARTICLE_SCALES, oldest first) and `count` particles at
    random reference positions with random reference velocities, last
    updated two time steps ago (particles.c is called every 2nd step)."""
    L = np.asarray(hi, float) - np.asarray(lo, float)
    X = 2 * np.pi * (mesh.x - lo[0]) / L[0]
    Y = 2 * np.pi * (mesh.y - lo[1]) / L[1]
    Zc = 2 * np.pi * (mesh.z - lo[2]) / L[2]
    u = np.sin(X) * np.cos(Y) * np.cos(Zc)
    v = -np.cos(X) * np.sin(Y) * np.cos(Zc)
    w = 0.5 * np.sin(X + Zc)
    vel = np.concatenate([u.ravel(), v.ravel(), w.ravel()])
    rng = np.random.default_rng(seed)
    rec = np.zeros((count, 9))
    rec[:, 0] = rng.integers(0, mesh.nel, count)
    rec[:, 1:4] = rng.uniform(-0.95, 0.95, (count, 3))
    rec[:, 4:7] = 1.5 * rng.standard_normal((count, 3))
    rec[:, 7] = PARTICLE_T - 2 * PARTICLE_DT
    rec[:, 8] = np.exp(rng.uniform(np.log(0.1), 0.0, count))
    return vel, rec


def write_mesh_bin(path, mesh):
    svvn = (2 * mesh.n) // 3
    with open(path, "wb") as fh:
        fh.write(np.array([mesh.n, mesh.nel], np.int32).tobytes())
        fh.write(np.array([mesh.dt, mesh.visc], np.float64).tobytes())
        fh.write(np.array([svvn], np.int32).tobytes())
        fh.write(np.array([0.0], np.float64).tobytes())
        for a in (mesh.neigh, mesh.edge_neigh, mesh.edge_size,
                  mesh.corn_neigh, mesh.corn_size):
            fh.write(np.ascontiguousarray(a, np.int32).tobytes())
        for a in (mesh.x, mesh.y, mesh.z):
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())


def rdb(d, name):
    return np.fromfile(os.path.join(d, name), dtype=np.float64)


def nmg_levels(n):
    out = []
    for i in range(10):
        v = max(2, (n - 1) // (2 ** i) + 1)
        out.append(v)
        if v == 2:
            break
    return out


def run_case(c):
    mesh = box_mesh(c["n"], c["m"], lo=c["lo"], hi=c["hi"], bc=c["bc"],
                    deform=c["deform"], amp=c["amp"], dt=c["dt"],
                    visc=c["visc"])
    rng = np.random.default_rng(c["seed"])
    nt = mesh.ntot
    levels = nmg_levels(mesh.n)
    u = rng.standard_normal(nt)
    nc = levels[1] if len(levels) > 1 else mesh.n
    v1 = rng.standard_normal(mesh.nel * nc ** 3)
    f_nodal = smooth_source(mesh.x, mesh.y, mesh.z, c["seed"] + 1000,
                            c["lo"], c["hi"])
    b = weak_rhs(mesh, f_nodal)
    with tempfile.TemporaryDirectory() as d:
        write_mesh_bin(os.path.join(d, "mesh.bin"), mesh)
        u.tofile(os.path.join(d, "in_u.bin"))
        b.tofile(os.path.join(d, "in_f.bin"))
        v1.tofile(os.path.join(d, "in_v1.bin"))
        pin = {}
        if c.get("particles"):
            vel, rec = particle_inputs(mesh, c["lo"], c["hi"], c["particles"],
                                       c["seed"] + 2000)
            vel.tofile(os.path.join(d, "in_vel.bin"))
            with open(os.path.join(d, "in_particles.bin"), "wb") as fh:
                fh.write(np.array([len(rec)], np.int32).tobytes())
                fh.write(np.array([PARTICLE_T, PARTICLE_DT,
                                   *PARTICLE_SCALES]).tobytes())
                fh.write(rec.tobytes())
            pin = dict(in_vel=vel, in_particles=rec,
                       particle_hdr=np.array([PARTICLE_T, PARTICLE_DT,
                                              *PARTICLE_SCALES]))
        env = dict(os.environ, MKL_CBWR="COMPATIBLE", MKL_NUM_THREADS="1",
                   MKL_THREADING_LAYER="SEQUENTIAL", OMP_NUM_THREADS="1")
        subprocess.run([HARNESS, "mesh", d], check=True, env=env)
        out = dict(
            n=np.int32(mesh.n), x=mesh.x.ravel(), y=mesh.y.ravel(),
            z=mesh.z.ravel(), neigh=mesh.neigh, edge_neigh=mesh.edge_neigh,
            edge_size=mesh.edge_size, corn_neigh=mesh.corn_neigh,
            corn_size=mesh.corn_size, dt=np.float64(mesh.dt),
            visc=np.float64(mesh.visc), in_u=u, in_f=b, in_v1=v1,
            levels=np.array(levels, np.int32), **pin,
        )
        for name in sorted(os.listdir(d)):
            if name.startswith("ref_"):
                out[name[:-4]] = rdb(d, name)
    sc = out["ref_scalars"]
    path = os.path.join(HERE, c["name"] + ".npz")
    np.savez_compressed(path, **out)
    return path, dict(
        n=c["n"], m=list(c["m"]), seed=c["seed"], levels=levels,
        eps=float(sc[0]), cg_helm_iters=int(sc[1]), fcg_pressure_iters=int(sc[3]),
        fcg_dirichlet_iters=int(sc[5]), bnorm=float(sc[7]))


def basis_fixture():
    path = os.path.join(HERE, "basis.npz")
    with tempfile.TemporaryDirectory() as d:
        raw = os.path.join(d, "basis.bin")
        subprocess.run([HARNESS, "basis", raw], check=True)
        a = np.fromfile(raw, dtype=np.float64)
    out, off = {}, 0
    for n in range(2, 17):
        for key, cnt in (("xi", n), ("rho", n), ("d", n * n), ("dsvv", n * n)):
            out[f"{key}_{n}"] = a[off:off + cnt]
            off += cnt
    for nf in range(3, 17):
        nc = max(2, (nf - 1) // 2 + 1)
        out[f"jcf_{nf}"] = a[off:off + nf * nc]
        off += nf * nc
        out[f"jfc_{nf}"] = a[off:off + nf * nc]
        off += nf * nc
    assert off == a.size
    np.savez_compressed(path, **out)
    return path


def example_fixture():
    """Connectivity and vertices of the reference's examples/grid.h5,
    dumped with h5dump (binary, memory layout)."""
    path = os.path.join(HERE, "example_grid.npz")
    with tempfile.TemporaryDirectory() as d:
        raw = os.path.join(d, "el.bin")
        subprocess.run([H5DUMP, "-b", "MEMORY", "-d", "/Element array", "-o", raw,
                        EXAMPLE_GRID], check=True, stdout=subprocess.DEVNULL)
        rec = np.dtype([("Neighbours", "<i4", (6,)),
                        ("Neighbour rotation", "<i4", (6,)),
                        ("Corner size", "<i4", (8,)),
                        ("Corner neighbours", "<i4", (8,)),
                        ("Edge size", "<i4", (12,)),
                        ("Edge neighbours", "<i4", (12,)),
                        ("X", "<f8", (8,)), ("Y", "<f8", (8,)),
                        ("Z", "<f8", (8,))])
        a = np.fromfile(raw, dtype=rec)
    np.savez_compressed(
        path, neigh=a["Neighbours"], rot=a["Neighbour rotation"],
        corn_size=a["Corner size"], corn_neigh=a["Corner neighbours"],
        edge_size=a["Edge size"], edge_neigh=a["Edge neighbours"],
        x=a["X"], y=a["Y"], z=a["Z"])
    return path


def sha(p):
    with open(p, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference first: make -C oracle ref")
    manifest = {"generator": "tests/golden/make_golden.py",
                "reference": "SSENSS 1.0.1 src/, compiled by oracle/Makefile",
                "blas": "MKL 2021.4.0 sequential, MKL_CBWR=COMPATIBLE",
                "files": {}}
    p = basis_fixture()
    manifest["files"][os.path.basename(p)] = {"sha256": sha(p)}
    if os.path.exists(EXAMPLE_GRID) and os.path.exists(H5DUMP):
        p = example_fixture()
        manifest["files"][os.path.basename(p)] = {"sha256": sha(p)}
    for c in CASES:
        p, info = run_case(c)
        info["sha256"] = sha(p)
        manifest["files"][os.path.basename(p)] = info
        print(c["name"], info, flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
