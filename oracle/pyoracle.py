"""ctypes binding of the oracle (oracle/libssx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker (or the timed CPU
baseline), never part of the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRESSURE, X_VELOCITY = 0, 1
OP_POISSON, OP_HELMHOLTZ = 0, 1
PC_JACOBI, PC_MG = 0, 1
# orc_level_data selectors
G0, DETJ, DIAG_P, DIAG_H, SCALE, MASK, INVMASS, X, Y, Z, DMAT, DSVV, PROL = (
    0, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17)


class _Mesh(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("nel", C.c_int),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("z", C.POINTER(C.c_double)),
        ("neigh", C.POINTER(C.c_int)), ("edge_neigh", C.POINTER(C.c_int)),
        ("edge_size", C.POINTER(C.c_int)), ("corn_neigh", C.POINTER(C.c_int)),
        ("corn_size", C.POINTER(C.c_int)),
        ("dt", C.c_double), ("visc", C.c_double),
        ("svvn", C.c_int), ("svva", C.c_double),
    ]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "libssx_oracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    i, d = C.c_int, C.c_double
    sig = {
        "orc_gll_basis": (None, [i, i, d, dp, dp, dp, dp]),
        "orc_interp_matrix": (None, [i, i, dp]),
        "orc_create": (vp, [C.POINTER(_Mesh)]),
        "orc_destroy": (None, [vp]),
        "orc_nlevels": (i, [vp]),
        "orc_level_n": (i, [vp, i]),
        "orc_nel": (i, [vp]),
        "orc_lambda": (d, [vp]),
        "orc_level_data": (dp, [vp, i, i]),
        "orc_ax": (None, [vp, i, i, dp, dp]),
        "orc_opeval": (None, [vp, i, i, i, dp, dp]),
        "orc_dss": (None, [vp, i, dp]),
        "orc_dss_average": (None, [vp, i, dp]),
        "orc_dot": (d, [vp, i, dp, dp]),
        "orc_mask": (None, [vp, i, i, dp]),
        "orc_subtract_nullspace": (None, [vp, i, i, dp]),
        "orc_restrict": (None, [vp, i, dp, dp]),
        "orc_prolong": (None, [vp, i, dp, dp]),
        "orc_cg": (d, [vp, i, i, i, dp, dp, d, i, ip]),
        "orc_fcg": (d, [vp, i, i, i, dp, dp, d, i, ip]),
        "orc_mg_precon": (None, [vp, i, i, d, dp, dp]),
        "orc_set_diag_choice": (None, [vp, i]),
        "orc_counters": (None, [vp, C.POINTER(C.c_longlong)]),
        "orc_set_threads": (None, [i]),
        "orc_get_threads": (i, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def set_threads(n):
    load().orc_set_threads(int(n))


def get_threads():
    return load().orc_get_threads()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gll_basis(n, svvn=None, svva=0.0):
    L = load()
    svvn = (2 * n) // 3 if svvn is None else svvn
    xi, rho = np.zeros(n), np.zeros(n)
    d, ds = np.zeros((n, n)), np.zeros((n, n))
    L.orc_gll_basis(n, svvn, svva, _dp(xi), _dp(rho), _dp(d), _dp(ds))
    return xi, rho, d, ds


def interp_matrix(nfrom, nto):
    j = np.zeros((nfrom, nto))
    load().orc_interp_matrix(nfrom, nto, _dp(j))
    return j


class Oracle:
    """CPU restatement of the reference path on one mesh."""

    def __init__(self, mesh):
        L = load()
        self.L = L
        self.keep = {}
        for a in ("x", "y", "z"):
            self.keep[a] = np.ascontiguousarray(getattr(mesh, a), np.float64)
        for a in ("neigh", "edge_neigh", "edge_size", "corn_neigh",
                  "corn_size"):
            self.keep[a] = np.ascontiguousarray(getattr(mesh, a), np.int32)
        ip = C.POINTER(C.c_int)
        svvn = mesh.svvn if getattr(mesh, "svvn", -1) >= 0 else (2 * mesh.n) // 3
        self.ms = _Mesh(mesh.n, self.keep["x"].shape[0],
                        _dp(self.keep["x"]), _dp(self.keep["y"]),
                        _dp(self.keep["z"]),
                        *(self.keep[a].ctypes.data_as(ip) for a in
                          ("neigh", "edge_neigh", "edge_size", "corn_neigh",
                           "corn_size")),
                        float(mesh.dt), float(mesh.visc), int(svvn),
                        float(getattr(mesh, "svva", 0.0)))
        self.h = L.orc_create(C.byref(self.ms))
        if not self.h:
            raise RuntimeError("orc_create failed (unsupported mesh)")
        self.nel = mesh.x.shape[0]

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nlevels(self):
        return self.L.orc_nlevels(self.h)

    def level_n(self, l):
        return self.L.orc_level_n(self.h, l)

    def ntot(self, l=0):
        return self.nel * self.level_n(l) ** 3

    @property
    def lam(self):
        return self.L.orc_lambda(self.h)

    def data(self, l, which):
        p = self.L.orc_level_data(self.h, l, which)
        if not p:
            return None
        n = self.level_n(l)
        if which in (DMAT, DSVV):
            cnt = n * n
        elif which == PROL:
            cnt = n * self.level_n(l + 1)
        else:
            cnt = self.ntot(l)
        return np.ctypeslib.as_array(p, shape=(cnt,)).copy()

    def ax(self, l, op, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        self.L.orc_ax(self.h, l, op, _dp(u), _dp(out))
        return out

    def opeval(self, l, op, var, p):
        p = np.ascontiguousarray(p, np.float64)
        out = np.zeros_like(p)
        self.L.orc_opeval(self.h, l, op, var, _dp(p), _dp(out))
        return out

    def dss(self, l, u):
        u = np.array(u, np.float64, copy=True)
        self.L.orc_dss(self.h, l, _dp(u))
        return u

    def dss_average(self, l, u):
        u = np.array(u, np.float64, copy=True)
        self.L.orc_dss_average(self.h, l, _dp(u))
        return u

    def dot(self, l, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        return self.L.orc_dot(self.h, l, _dp(x), _dp(y))

    def mask(self, l, var, u):
        u = np.array(u, np.float64, copy=True)
        self.L.orc_mask(self.h, l, var, _dp(u))
        return u

    def nullspace(self, l, var, u):
        u = np.array(u, np.float64, copy=True)
        self.L.orc_subtract_nullspace(self.h, l, var, _dp(u))
        return u

    def restrict(self, l, fine):
        fine = np.ascontiguousarray(fine, np.float64)
        out = np.zeros(self.ntot(l))
        self.L.orc_restrict(self.h, l, _dp(fine), _dp(out))
        return out

    def prolong(self, l, coarse):
        coarse = np.ascontiguousarray(coarse, np.float64)
        out = np.zeros(self.ntot(l))
        self.L.orc_prolong(self.h, l, _dp(coarse), _dp(out))
        return out

    def set_diag_choice(self, c):
        self.L.orc_set_diag_choice(self.h, c)

    def cg(self, l, op, var, f, u, eps, max_iter):
        f = np.array(f, np.float64, copy=True)
        u = np.array(u, np.float64, copy=True)
        it = C.c_int()
        err = self.L.orc_cg(self.h, l, op, var, _dp(f), _dp(u), eps, max_iter,
                            C.byref(it))
        return u, it.value, err, f

    def fcg(self, op, var, precon, f, u, eps, max_iter):
        f = np.array(f, np.float64, copy=True)
        u = np.array(u, np.float64, copy=True)
        it = C.c_int()
        err = self.L.orc_fcg(self.h, op, var, precon, _dp(f), _dp(u), eps,
                             max_iter, C.byref(it))
        return u, it.value, err, f

    def mg_precon(self, op, var, eps, x):
        x = np.array(x, np.float64, copy=True)
        px = np.zeros_like(x)
        self.L.orc_mg_precon(self.h, op, var, eps, _dp(x), _dp(px))
        return px

    def counters(self):
        a = (C.c_longlong * 4)()
        self.L.orc_counters(self.h, a)
        return tuple(a)


class Particle(C.Structure):
    """orc_particle (struct particle of the reference, single rank)"""
    _fields_ = [("id", C.c_long), ("el", C.c_int), ("old_el", C.c_int),
                ("status", C.c_int), ("x", C.c_double), ("y", C.c_double),
                ("z", C.c_double), ("u", C.c_double), ("v", C.c_double),
                ("w", C.c_double), ("t", C.c_double), ("tau", C.c_double),
                ("reinit", C.c_int)]


def _particle_sigs(L):
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.orc_interp_point.restype = C.c_double
    L.orc_interp_point.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double,
                                   C.c_double]
    L.orc_particles_set_velocity.restype = None
    L.orc_particles_set_velocity.argtypes = [vp, dp, dp, dp]
    L.orc_particles_step.restype = C.c_int
    L.orc_particles_step.argtypes = [vp, C.POINTER(Particle), C.c_long,
                                     C.c_double, C.c_double]
    L.orc_particle_position.restype = None
    L.orc_particle_position.argtypes = [vp, C.POINTER(Particle), dp]
    L.orc_tracers_step.restype = C.c_int
    L.orc_tracers_step.argtypes = [vp, C.POINTER(Particle), C.c_long,
                                   C.c_double, C.c_double]


def interp_point(n, xi, f, x, y, z):
    L = load()
    _particle_sigs(L)
    xi = np.ascontiguousarray(xi, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    return L.orc_interp_point(n, _dp(xi), _dp(f), x, y, z)


def particles_from_records(rec):
    """rec: (n, 9) el, x, y, z, u, v, w, t, tau"""
    arr = (Particle * len(rec))()
    for i, r in enumerate(rec):
        p = arr[i]
        p.id, p.el, p.old_el, p.status = i, int(r[0]), int(r[0]), 0
        p.x, p.y, p.z, p.u, p.v, p.w, p.t, p.tau = map(float, r[1:9])
        p.reinit = 0
    return arr


def particles_to_records(arr):
    """(n, 8): el, x, y, z, u, v, w, t  and the re-initialised flags"""
    out = np.array([[p.el, p.x, p.y, p.z, p.u, p.v, p.w, p.t] for p in arr])
    return out, np.array([p.reinit for p in arr])


def _oracle_particles(self, vel, scales, rec, t_now, dt):
    """Three set_velocity calls (oldest first), then one step."""
    L = self.L
    _particle_sigs(L)
    nt = self.ntot()
    for s in scales:
        u = np.ascontiguousarray(s * vel[:nt])
        v = np.ascontiguousarray(s * vel[nt:2 * nt])
        w = np.ascontiguousarray(s * vel[2 * nt:])
        L.orc_particles_set_velocity(self.h, _dp(u), _dp(v), _dp(w))
    arr = particles_from_records(rec)
    passes = L.orc_particles_step(self.h, arr, len(rec), t_now, dt)
    out, re = particles_to_records(arr)
    return out, re, passes


Oracle.particles = _oracle_particles


def _oracle_tracers(self, vel, scales, rec, t_now, dt, steps=1):
    """Three set_velocity calls (oldest first), then `steps` RK2 tracer
    steps of dt ending at t_now (orc_tracers_step).  rec as particles();
    returns (n, 8) records, reinit flags, the largest pass count and the
    physical positions (n, 3)."""
    L = self.L
    _particle_sigs(L)
    nt = self.ntot()
    for s in scales:
        u = np.ascontiguousarray(s * vel[:nt])
        v = np.ascontiguousarray(s * vel[nt:2 * nt])
        w = np.ascontiguousarray(s * vel[2 * nt:])
        L.orc_particles_set_velocity(self.h, _dp(u), _dp(v), _dp(w))
    arr = particles_from_records(rec)
    passes = 0
    for k in range(steps):
        tk = t_now - (steps - 1 - k) * dt
        passes = max(passes, L.orc_tracers_step(self.h, arr, len(rec), tk,
                                                dt))
    out, re = particles_to_records(arr)
    xyz = np.zeros((len(rec), 3))
    for i in range(len(rec)):
        L.orc_particle_position(self.h, C.byref(arr[i]),
                                xyz[i].ctypes.data_as(C.POINTER(C.c_double)))
    return out, re, passes, xyz


Oracle.tracers = _oracle_tracers


def _oracle_positions(self, out):
    """physical coordinates (n, 3) of records (n, >=4): el, x, y, z"""
    L = self.L
    _particle_sigs(L)
    xyz = np.zeros((len(out), 3))
    p = Particle()
    for i, r in enumerate(out):
        p.el, p.x, p.y, p.z = int(r[0]), float(r[1]), float(r[2]), float(r[3])
        L.orc_particle_position(self.h, C.byref(p),
                                xyz[i].ctypes.data_as(C.POINTER(C.c_double)))
    return xyz


Oracle.positions = _oracle_positions
