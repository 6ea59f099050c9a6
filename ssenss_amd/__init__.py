"""ssenss_amd -- MI355X-native elliptic-solve engine for SSENSS.

The engine is the C-ABI library ``libssx.so`` (C host code + gfx950 HIP
kernels, declared in ``include/ssx.h``).  This package only loads it with
ctypes and gives tests and benchmarks a thin Python handle; no numerics run
in Python and there is no fallback path: if the library is missing, every
entry point raises.
"""
from .native import (  # noqa: F401
    SsxError, lib, load, library_path, Stats,
    PRESSURE, X_VELOCITY, Y_VELOCITY, Z_VELOCITY,
    OP_POISSON, OP_HELMHOLTZ, PC_JACOBI, PC_MG, OPT_FUSED_CG,
    OPT_SMOOTHER_PRECISION, OPT_PERSISTENT_CG, OPT_LP_COARSE_RTOL,
    OPT_FUSED_REDUCE, OPT_COARSE_ASSEMBLED, OPT_GATHER_DSS,
    OPT_PERSIST_ABORT_TEST,
    OPT_PARTICLE_WALK,
    OPT_COARSE_DIRECT, OPT_COARSE_SPARSE,
    PREC_FP64, PREC_FP32, PREC_BF16, NOALLOC_MAX_PARTICLES,
    AX_NVARIANTS,
    DATA_GFAC, DATA_MASS, DATA_DIAG_P, DATA_DIAG_H, DATA_SCALE, DATA_MASK,
    DATA_INVMASS, DATA_X, DATA_Y, DATA_Z, EXPORTED_SYMBOLS,
    NS_U, NS_V, NS_W, NS_P,
)
from .mesh import Mesh, box_mesh  # noqa: F401
from .engine import Batch, Context, NavierStokes  # noqa: F401
