"""ctypes binding of libssx_io.so (include/ssx_io.h): SSENSS's grid.h5 /
init.h5 / vel_%08d.h5 files and setup input."""
import ctypes as C
import os

import numpy as np

from .native import ERRORS, SsxError, load

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class Setup(C.Structure):
    _fields_ = [("dt", C.c_double), ("tmax", C.c_double),
                ("visc", C.c_double), ("dens", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("n", C.c_int), ("write_every", C.c_int), ("test", C.c_int),
                ("datdir", C.c_char * 100), ("gridfile", C.c_char * 100),
                ("initfile", C.c_char * 100)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = v.decode() if isinstance(v, bytes) else v
        return out


def library_path():
    return os.environ.get("SSX_IO_LIB", os.path.join(_HERE, "libssx_io.so"))


def lib():
    global _lib
    if _lib is None:
        load()                      # libssx.so first (GLL basis)
        L = C.CDLL(library_path())
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        s, i = C.c_char_p, C.c_int
        L.ssx_io_last_error.restype = C.c_char_p
        sig = {
            "ssx_io_file_params": [s, ip, ip],
            "ssx_io_read_grid": [s, i, i, dp, dp, dp, ip, ip, ip, ip, ip, ip],
            "ssx_io_write_grid": [s, i, i, dp, dp, dp, ip, ip, ip, ip, ip,
                                  ip],
            "ssx_io_read_velocity": [s, i, i, dp, dp, dp],
            "ssx_io_write_velocity": [s, i, i, dp, dp, dp],
            "ssx_io_read_setup": [s, C.POINTER(Setup)],
            "ssx_io_parse_args": [i, C.POINTER(C.c_char_p),
                                  C.POINTER(Setup)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        L.ssx_io_setup_defaults.restype = None
        L.ssx_io_setup_defaults.argtypes = [C.POINTER(Setup)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().ssx_io_last_error()
        raise SsxError(rc, msg.decode() if msg else ERRORS.get(rc, ""))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def file_params(path):
    n, nel = C.c_int(), C.c_int()
    _check(lib().ssx_io_file_params(path.encode(), C.byref(n), C.byref(nel)))
    return n.value, nel.value


def read_grid(path, n_run=None):
    """dict of x, y, z (nel, n_run^3) and int connectivity arrays"""
    n, nel = file_params(path)
    n_run = n if n_run is None else n_run
    r3 = n_run ** 3
    out = {k: np.zeros((nel, r3)) for k in ("x", "y", "z")}
    shapes = {"neigh": 6, "rot": 6, "edge_neigh": 12, "edge_size": 12,
              "corn_neigh": 8, "corn_size": 8}
    out.update({k: np.zeros((nel, m), np.int32) for k, m in shapes.items()})
    _check(lib().ssx_io_read_grid(
        path.encode(), n_run, nel, _dp(out["x"]), _dp(out["y"]),
        _dp(out["z"]), _ip(out["neigh"]), _ip(out["rot"]),
        _ip(out["edge_neigh"]), _ip(out["edge_size"]), _ip(out["corn_neigh"]),
        _ip(out["corn_size"])))
    out["n"] = n_run
    return out


def write_grid(path, mesh):
    n, nel = mesh.n, mesh.nel
    arr = lambda a, t: np.ascontiguousarray(a, t).reshape(nel, -1)
    xs = [arr(getattr(mesh, k), np.float64) for k in ("x", "y", "z")]
    ints = [arr(getattr(mesh, k), np.int32) for k in
            ("neigh", "edge_neigh", "edge_size", "corn_neigh", "corn_size")]
    rot = np.zeros((nel, 6), np.int32)
    _check(lib().ssx_io_write_grid(path.encode(), n, nel, *(_dp(a) for a in xs),
                                   _ip(ints[0]), _ip(rot), *(_ip(a) for a in
                                                            ints[1:])))


def read_velocity(path, nel, n_run=None):
    n, neld = file_params(path)
    n_run = n if n_run is None else n_run
    u, v, w = (np.zeros((nel, n_run ** 3)) for _ in range(3))
    _check(lib().ssx_io_read_velocity(path.encode(), n_run, nel, _dp(u),
                                      _dp(v), _dp(w)))
    return u, v, w


def write_velocity(path, n, u, v, w):
    u, v, w = (np.ascontiguousarray(a, np.float64) for a in (u, v, w))
    nel = u.size // n ** 3
    _check(lib().ssx_io_write_velocity(path.encode(), n, nel, _dp(u), _dp(v),
                                       _dp(w)))


def read_setup(path, argv=None):
    """the setup file's parameters block over the defaults, then the
    command-line overrides of argv (argv[0] is the program name)"""
    s = Setup()
    lib().ssx_io_setup_defaults(C.byref(s))
    _check(lib().ssx_io_read_setup(path.encode(), C.byref(s)))
    if argv:
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        _check(lib().ssx_io_parse_args(len(argv), arr, C.byref(s)))
    return s.as_dict()
