"""ctypes binding of libssx.so (include/ssx.h)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

PRESSURE, X_VELOCITY, Y_VELOCITY, Z_VELOCITY = 0, 1, 2, 3
OP_POISSON, OP_HELMHOLTZ = 0, 1
PC_JACOBI, PC_MG = 0, 1
OPT_FUSED_CG = 1
OPT_SMOOTHER_PRECISION = 2
OPT_PERSISTENT_CG = 3
OPT_LP_COARSE_RTOL = 4
OPT_FUSED_REDUCE = 5
OPT_COARSE_ASSEMBLED = 6
OPT_GATHER_DSS = 7
OPT_PERSIST_ABORT_TEST = 8
OPT_PARTICLE_WALK = 9
OPT_COARSE_DIRECT = 10
OPT_COARSE_SPARSE = 11
AX_NVARIANTS = 8
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
NOALLOC_MAX_PARTICLES = 131072
(DATA_GFAC, DATA_MASS, DATA_DIAG_P, DATA_DIAG_H, DATA_SCALE, DATA_MASK,
 DATA_INVMASS, DATA_X, DATA_Y, DATA_Z) = range(10)

ERRORS = {1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ENOTCONV", 5: "EMESH",
          6: "ENODEV", 7: "ENUMERIC"}

# every function declared in include/ssx.h
EXPORTED_SYMBOLS = [
    "ssx_last_error", "ssx_version", "ssx_mesh_box", "ssx_dss_host",
    "ssx_dss_gather_host",
    "ssx_geometry_host", "ssx_create",
    "ssx_destroy", "ssx_set_stream", "ssx_own_stream", "ssx_get_stream", "ssx_synchronize",
    "ssx_nlevels", "ssx_level_n", "ssx_nel", "ssx_level_ntot", "ssx_lambda",
    "ssx_level_data", "ssx_dss_groups", "ssx_set_diag_choice",
    "ssx_set_option", "ssx_ax",
    "ssx_opeval", "ssx_dss", "ssx_dss_average", "ssx_dot", "ssx_mask",
    "ssx_nullspace", "ssx_restrict", "ssx_prolong", "ssx_jacobi", "ssx_cg",
    "ssx_fcg", "ssx_mg_precon", "ssx_solve_poisson_cg",
    "ssx_solve_helmholtz_cg", "ssx_particles_set_velocity",
    "ssx_particles_step", "ssx_particles_set_velocity_host",
    "ssx_particles_step_host", "ssx_tracers_step", "ssx_tracers_step_host",
    "ssx_autotune", "ssx_ax_variant",
    "ssx_set_ax_variant", "ssx_batch_create", "ssx_batch_destroy",
    "ssx_batch_set_stream", "ssx_batch_cg",
    "ssx_particles_step_host_noalloc", "ssx_batch_solve",
    "ssx_ns_create", "ssx_ns_destroy", "ssx_ns_init", "ssx_ns_step",
    "ssx_ns_field", "ssx_ns_read", "ssx_ns_time",
    "ssx_particles_pass_counts",
]


NS_U, NS_V, NS_W, NS_P = 0, 1, 2, 3


class NSParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("visc", C.c_double),
                ("dens", C.c_double), ("t", C.c_double),
                ("g", C.c_double * 3), ("batch_velocity", C.c_int)]


class SsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ssx error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class MeshStruct(C.Structure):
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


class ParticlesStruct(C.Structure):
    _fields_ = [("n", C.c_long)] + [(f, C.c_void_p) for f in (
        "el", "old_el", "x", "y", "z", "u", "v", "w", "t", "tau", "reinit")]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("converged", C.c_int),
        ("residual", C.c_double), ("fine_ax", C.c_longlong),
        ("coarse_iters", C.c_longlong), ("smooth_iters", C.c_longlong),
        ("vcycles", C.c_longlong), ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def library_path():
    return os.path.join(_HERE, "libssx.so")


_lib = None


def load():
    """Load libssx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `make -C ssenss_amd/csrc` or "
            "__graft_entry__.build(); there is no fallback path")
    # PyTorch-ROCm ships its own HIP runtime libraries.  Loaded after
    # libssx.so pulled in the system ROCm ones, torch then finds no GPU
    # ("No HIP GPUs are available"), so torch, when installed, is imported
    # first; the device tensors it allocates are what the caller passes.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    i, d, sz = C.c_int, C.c_double, C.c_size_t
    sig = {
        "ssx_last_error": (C.c_char_p, []),
        "ssx_version": (C.c_char_p, []),
        "ssx_mesh_box": (i, [i, i, i, i, dp, dp, ip, i, d, dp, dp, dp, ip, ip,
                             ip, ip, ip]),
        "ssx_dss_host": (i, [C.POINTER(MeshStruct), i, dp]),
        "ssx_dss_gather_host": (i, [C.POINTER(MeshStruct), i, dp]),
        "ssx_geometry_host": (i, [C.POINTER(MeshStruct), dp, dp]),
        "ssx_create": (i, [C.POINTER(vp), C.POINTER(MeshStruct)]),
        "ssx_destroy": (None, [vp]),
        "ssx_set_stream": (i, [vp, vp]),
        "ssx_get_stream": (vp, [vp]),
        "ssx_own_stream": (vp, [vp]),
        "ssx_synchronize": (i, [vp]),
        "ssx_nlevels": (i, [vp]),
        "ssx_level_n": (i, [vp, i]),
        "ssx_nel": (i, [vp]),
        "ssx_level_ntot": (sz, [vp, i]),
        "ssx_lambda": (d, [vp]),
        "ssx_level_data": (i, [vp, i, i, dp]),
        "ssx_dss_groups": (i, [vp, i, C.POINTER(C.c_longlong)]),
        "ssx_set_diag_choice": (i, [vp, i]),
        "ssx_set_option": (i, [vp, i, i]),
        "ssx_ax": (i, [vp, i, i, vp, vp]),
        "ssx_opeval": (i, [vp, i, i, i, vp, vp]),
        "ssx_dss": (i, [vp, i, vp]),
        "ssx_dss_average": (i, [vp, i, vp]),
        "ssx_dot": (i, [vp, i, vp, vp, dp]),
        "ssx_mask": (i, [vp, i, i, vp]),
        "ssx_nullspace": (i, [vp, i, i, vp]),
        "ssx_restrict": (i, [vp, i, vp, vp]),
        "ssx_prolong": (i, [vp, i, vp, vp]),
        "ssx_jacobi": (i, [vp, i, i, vp, vp]),
        "ssx_cg": (i, [vp, i, i, i, vp, vp, d, i, C.POINTER(Stats)]),
        "ssx_fcg": (i, [vp, i, i, i, vp, vp, d, i, C.POINTER(Stats)]),
        "ssx_mg_precon": (i, [vp, i, i, d, vp, vp]),
        "ssx_solve_poisson_cg": (i, [vp, dp, dp, i, d, C.POINTER(Stats)]),
        "ssx_solve_helmholtz_cg": (i, [vp, dp, dp, i, d, C.POINTER(Stats)]),
        "ssx_particles_set_velocity": (i, [vp, vp, vp, vp]),
        "ssx_particles_set_velocity_host": (i, [vp, dp, dp, dp]),
        "ssx_autotune": (i, [vp, i]),
        "ssx_batch_create": (i, [C.POINTER(vp), C.POINTER(vp), ip, ip, i]),
        "ssx_batch_destroy": (None, [vp]),
        "ssx_batch_set_stream": (i, [vp, vp]),
        "ssx_batch_cg": (i, [vp, C.POINTER(vp), C.POINTER(vp), dp, i, i,
                             C.POINTER(Stats)]),
        "ssx_batch_solve": (i, [vp, ip, C.POINTER(vp), C.POINTER(vp), dp, i,
                                i, C.POINTER(Stats)]),
        "ssx_ns_create": (i, [C.POINTER(vp), vp, C.POINTER(NSParams)]),
        "ssx_ns_destroy": (None, [vp]),
        "ssx_ns_init": (i, [vp, C.POINTER(vp)]),
        "ssx_ns_step": (i, [vp, C.POINTER(Stats)]),
        "ssx_ns_field": (vp, [vp, i, i]),
        "ssx_ns_read": (i, [vp, i, i, vp]),
        "ssx_ns_time": (d, [vp]),
        "ssx_ax_variant": (i, [vp, i, i, dp]),
        "ssx_set_ax_variant": (i, [vp, i, i, i]),
        "ssx_particles_step_host": (i, [vp, C.POINTER(ParticlesStruct), d, d,
                                        C.POINTER(C.c_int)]),
        "ssx_particles_step": (i, [vp, C.POINTER(ParticlesStruct), d, d,
                                   C.POINTER(C.c_int)]),
        "ssx_tracers_step": (i, [vp, C.POINTER(ParticlesStruct), d, d,
                                 C.POINTER(C.c_int)]),
        "ssx_tracers_step_host": (i, [vp, C.POINTER(ParticlesStruct), d, d,
                                      C.POINTER(C.c_int)]),
        "ssx_particles_step_host_noalloc": (
            i, [vp, C.POINTER(ParticlesStruct), d, d, C.POINTER(C.c_int)]),
        "ssx_particles_pass_counts": (i, [vp, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def lib():
    return load()


def check(rc):
    if rc != 0:
        msg = load().ssx_last_error()
        raise SsxError(rc, msg.decode() if msg else "")
    return rc
