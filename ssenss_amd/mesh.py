"""Mesh container (numpy arrays in the grid.h5 field layout) and the
structured box generator of libssx (ssx_mesh_box)."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from .native import MeshStruct, check, load

NOSLIP, FREESLIP, PERIODIC = -1, -2, 0


@dataclass
class Mesh:
    n: int                    # GLL points per direction
    x: np.ndarray             # (nel, n, n, n) float64, [e][k][j][i]
    y: np.ndarray
    z: np.ndarray
    neigh: np.ndarray         # (nel, 6) int32, W E S N B T
    edge_neigh: np.ndarray    # (nel, 12)
    edge_size: np.ndarray     # (nel, 12)
    corn_neigh: np.ndarray    # (nel, 8)
    corn_size: np.ndarray     # (nel, 8)
    dt: float = 1.0
    visc: float = 1.0
    svvn: int = -1            # default 2n/3 (reference main.c:54)
    svva: float = 0.0         # reference main.c:55

    @property
    def nel(self):
        return self.x.shape[0]

    @property
    def ntot(self):
        return self.nel * self.n ** 3

    def struct(self):
        """MeshStruct view; the arrays must stay alive while it is used."""
        for a in ("x", "y", "z"):
            setattr(self, a, np.ascontiguousarray(getattr(self, a),
                                                  dtype=np.float64))
        for a in ("neigh", "edge_neigh", "edge_size", "corn_neigh",
                  "corn_size"):
            setattr(self, a, np.ascontiguousarray(getattr(self, a),
                                                  dtype=np.int32))
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        svvn = self.svvn if self.svvn >= 0 else (2 * self.n) // 3
        return MeshStruct(
            self.n, self.nel,
            self.x.ctypes.data_as(dp), self.y.ctypes.data_as(dp),
            self.z.ctypes.data_as(dp),
            self.neigh.ctypes.data_as(ip), self.edge_neigh.ctypes.data_as(ip),
            self.edge_size.ctypes.data_as(ip),
            self.corn_neigh.ctypes.data_as(ip),
            self.corn_size.ctypes.data_as(ip),
            float(self.dt), float(self.visc), int(svvn), float(self.svva))


def box_mesh(n, m, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             bc=(NOSLIP,) * 6, deform=0, amp=0.0, dt=1.0, visc=1.0):
    """Structured box of m=(mx,my,mz) hexahedra with n GLL points per
    direction.  deform 1: s = amp*sin(pi u)sin(pi v)sin(pi w) added to each
    coordinate (unit-cube coordinates u,v,w); deform 2: tanh stretching of y
    with factor amp."""
    if isinstance(m, int):
        m = (m, m, m)
    mx, my, mz = m
    nel = mx * my * mz
    L = load()
    x = np.zeros((nel, n, n, n))
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    ar = {k: np.zeros((nel, w), dtype=np.int32) for k, w in
          (("neigh", 6), ("edge_neigh", 12), ("edge_size", 12),
           ("corn_neigh", 8), ("corn_size", 8))}
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lo_a = (C.c_double * 3)(*lo)
    hi_a = (C.c_double * 3)(*hi)
    bc_a = (C.c_int * 6)(*bc)
    check(L.ssx_mesh_box(n, mx, my, mz, lo_a, hi_a, bc_a, deform, amp,
                         x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                         z.ctypes.data_as(dp),
                         *(ar[k].ctypes.data_as(ip) for k in
                           ("neigh", "edge_neigh", "edge_size", "corn_neigh",
                            "corn_size"))))
    return Mesh(n, x, y, z, ar["neigh"], ar["edge_neigh"], ar["edge_size"],
                ar["corn_neigh"], ar["corn_size"], dt=dt, visc=visc)
