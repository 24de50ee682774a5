"""ctypes binding of liborf_rcwa.so (include/orf_rcwa.h, include/orf_pageops.h).

The shared library is built in-tree by ``__graft_entry__.build()``
(opticrcwafdtd_amd/build_lib.sh).  There is no CPU fallback: if the library
is missing or no gfx950 device can be opened, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborf_rcwa.so")

ORF_OK, ORF_INVALID_ARG, ORF_INVALID_POL, ORF_LOSSY_INCIDENT, ORF_OOM, ORF_CANCELLED, \
    ORF_DEVICE_ERROR, ORF_SAMPLING = range(8)

# status code -> reference MATLAB error identifier
ERROR_IDS = {
    ORF_INVALID_ARG: "orf:invalidArgument",
    ORF_INVALID_POL: "rcwa:pol",
    ORF_LOSSY_INCIDENT: "rcwa:lossyIncident",
    ORF_OOM: "orf:outOfMemory",
    ORF_CANCELLED: "orf:cancelled",
    ORF_DEVICE_ERROR: "orf:device",
    ORF_SAMPLING: "convMatrix1D:sampling",
}


class orf_shape(C.Structure):
    _fields_ = [("type", C.c_int32), ("material", C.c_int32), ("cx", C.c_double), ("cy", C.c_double),
                ("a", C.c_double), ("b", C.c_double), ("vert0", C.c_int32), ("nvert", C.c_int32)]


class orf_problem(C.Structure):
    _fields_ = [
        ("Lx", C.c_double), ("Ly", C.c_double), ("Nx", C.c_int32), ("Ny", C.c_int32),
        ("Nsx", C.c_int32), ("Nsy", C.c_int32), ("nLambda", C.c_int32),
        ("lambda0", C.POINTER(C.c_double)), ("theta", C.c_double), ("phi", C.c_double),
        ("nPol", C.c_int32), ("polKind", C.POINTER(C.c_int32)), ("jones", C.POINTER(C.c_double)),
        ("nMaterials", C.c_int32), ("matEps", C.POINTER(C.c_double)),
        ("incMaterial", C.c_int32), ("subMaterial", C.c_int32),
        ("nStruct", C.c_int32), ("nLayers", C.c_int32),
        ("thickness", C.POINTER(C.c_double)), ("background", C.POINTER(C.c_int32)),
        ("shapeStart", C.POINTER(C.c_int32)), ("shapeCount", C.POINTER(C.c_int32)),
        ("nShapes", C.c_int32), ("shapes", C.POINTER(orf_shape)),
        ("nVertices", C.c_int32), ("vertices", C.POINTER(C.c_double)),
        ("factorization", C.c_int32), ("symmetry", C.c_int32), ("gapEps", C.c_double),
        ("precision", C.c_int32), ("batchSize", C.c_int32),
        ("structPeriod", C.POINTER(C.c_double)), ("nPages", C.c_int32),
        ("pageStruct", C.POINTER(C.c_int32)), ("pageLambda", C.POINTER(C.c_int32)),
        ("keepModes", C.c_int32),
    ]


class orf_result(C.Structure):
    _fields_ = [("m", C.POINTER(C.c_int32)), ("n", C.POINTER(C.c_int32)),
                ("R", C.POINTER(C.c_double)), ("T", C.POINTER(C.c_double)),
                ("Rtot", C.POINTER(C.c_double)), ("Ttot", C.POINTER(C.c_double)),
                ("r", C.POINTER(C.c_double)), ("t", C.POINTER(C.c_double))]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_char_p, C.c_void_p)

EXPORTS = [
    "orf_init", "orf_release", "orf_last_error", "orf_harmonic_count", "orf_rcwa2d",
    "orf_plan_create", "orf_plan_execute", "orf_plan_fetch", "orf_plan_destroy", "orf_plan_info",
    "orf_plan_stage_times", "orf_plan_enable_timing", "orf_plan_memory", "orf_plan_estimate",
    "orf_release_cached_memory", "orf_fields",
    "orf_pagemtimes", "orf_pagemtimes_f32", "orf_pagemldivide", "orf_pagemrdivide", "orf_pageeig",
    "orf_pagemldivide_noalloc", "orf_plan_enable_graphs",
    "orf_tune_set", "orf_tune_force", "orf_tune_clear", "orf_tune_load",
]
TUNING_TABLE = os.path.join(_HERE, "tuning_table.txt")

_lib = None
_ctx = None
_lock = threading.Lock()


def load_library():
    """Load liborf_rcwa.so and declare the prototypes (no device is touched)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.orf_init.argtypes = [C.c_int, P(C.c_void_p)]
    lib.orf_init.restype = C.c_int
    lib.orf_release.argtypes = [C.c_void_p]
    lib.orf_release.restype = None
    lib.orf_last_error.argtypes = []
    lib.orf_last_error.restype = C.c_char_p
    lib.orf_harmonic_count.argtypes = [P(orf_problem)]
    lib.orf_harmonic_count.restype = C.c_int
    lib.orf_rcwa2d.argtypes = [C.c_void_p, P(orf_problem), P(orf_result), PROGRESS_FN, C.c_void_p]
    lib.orf_rcwa2d.restype = C.c_int
    lib.orf_plan_create.argtypes = [C.c_void_p, P(orf_problem), P(C.c_void_p)]
    lib.orf_plan_create.restype = C.c_int
    lib.orf_plan_execute.argtypes = [C.c_void_p, PROGRESS_FN, C.c_void_p]
    lib.orf_plan_execute.restype = C.c_int
    lib.orf_plan_fetch.argtypes = [C.c_void_p, P(orf_result)]
    lib.orf_plan_fetch.restype = C.c_int
    lib.orf_plan_destroy.argtypes = [C.c_void_p]
    lib.orf_plan_destroy.restype = None
    lib.orf_plan_info.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_int32)]
    lib.orf_plan_info.restype = C.c_int
    lib.orf_plan_stage_times.argtypes = [C.c_void_p, C.c_int32, P(C.c_char_p), P(C.c_double),
                                         P(C.c_int32), P(C.c_int32)]
    lib.orf_plan_stage_times.restype = C.c_int
    lib.orf_plan_enable_timing.argtypes = [C.c_void_p, C.c_int32]
    lib.orf_plan_enable_timing.restype = None
    lib.orf_tune_set.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    lib.orf_tune_set.restype = C.c_int
    lib.orf_tune_force.argtypes = [C.c_int32, C.c_int32]
    lib.orf_tune_force.restype = C.c_int
    lib.orf_tune_clear.argtypes = []
    lib.orf_tune_clear.restype = None
    lib.orf_tune_load.argtypes = [C.c_char_p]
    lib.orf_tune_load.restype = C.c_int
    lib.orf_plan_enable_graphs.argtypes = [C.c_void_p, C.c_int32]
    lib.orf_plan_enable_graphs.restype = None
    lib.orf_plan_memory.argtypes = [C.c_void_p, P(C.c_int64)]
    lib.orf_plan_memory.restype = C.c_int
    lib.orf_plan_estimate.argtypes = [P(orf_problem), C.c_int32, P(C.c_int64)]
    lib.orf_plan_estimate.restype = C.c_int
    lib.orf_fields.argtypes = [C.c_void_p, P(orf_problem), C.c_int32, P(C.c_double), C.c_int32, P(C.c_double),
                               C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double),
                               P(orf_result)]
    lib.orf_fields.restype = C.c_int
    lib.orf_release_cached_memory.argtypes = []
    lib.orf_release_cached_memory.restype = None
    lib.orf_pagemtimes.argtypes = [C.c_void_p, C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32,
                                   P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, P(C.c_double),
                                   C.c_int32]
    lib.orf_pagemtimes.restype = C.c_int
    lib.orf_pagemtimes_f32.argtypes = lib.orf_pagemtimes.argtypes
    lib.orf_pagemtimes_f32.restype = C.c_int
    lib.orf_pagemldivide.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double), C.c_int64,
                                     P(C.c_double), P(C.c_double), C.c_int32]
    lib.orf_pagemldivide.restype = C.c_int
    lib.orf_pagemldivide_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                             C.c_void_p, C.c_int32, C.c_void_p]
    lib.orf_pagemldivide_noalloc.restype = C.c_int
    lib.orf_pagemrdivide.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double), P(C.c_double),
                                     C.c_int64, P(C.c_double), C.c_int32]
    lib.orf_pagemrdivide.restype = C.c_int
    lib.orf_pageeig.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double),
                                C.c_int32]
    lib.orf_pageeig.restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    return load_library().orf_last_error().decode(errors="replace")


def context():
    """The process-wide context on HIP device 0 (opened on first use)."""
    global _ctx
    with _lock:
        if _ctx is None:
            lib = load_library()
            h = C.c_void_p()
            rc = lib.orf_init(0, C.byref(h))
            if rc != ORF_OK:
                raise RuntimeError(f"orf_init failed ({rc}): {last_error()}")
            # the autotuned kernel choices (opticrcwafdtd_amd/autotune.py) when present
            if os.path.exists(TUNING_TABLE) and lib.orf_tune_load(TUNING_TABLE.encode()) != ORF_OK:
                raise RuntimeError(f"orf_tune_load failed: {last_error()}")
            _ctx = h
        return _ctx


def dptr(a):
    """double* view of a contiguous float64 / complex128 numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_double))
