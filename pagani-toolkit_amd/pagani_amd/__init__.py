"""Python host mirror of the libpagani C ABI (include/pagani.h).

Arrays may be numpy arrays (host) or torch CUDA tensors (device); the library
detects which.  The shared library must be built in-tree
(``python pagani-toolkit_amd/build.py``); importing this package fails loudly
when it is missing -- there is no CPU fallback.
"""
from .lib import (  # noqa: F401
    CorrInfo,
    CohortInfo,
    Context,
    Subject,
    TuneResult,
    PgError,
    LIB_PATH,
    PREC_FP32,
    PREC_BF16X3,
    THR_R,
    THR_SPARSITY,
    LAYOUT_TIME_MAJOR,
    LAYOUT_VOXEL_MAJOR,
    CP_BARRAT,
    CP_ONNELA,
    MOD_STRENGTH,
    MOD_BINARY_K,
    MOD_STOP_VECTOR,
    MOD_STOP_RAYLEIGH,
    library,
)
from . import synth  # noqa: F401
