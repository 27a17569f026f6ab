"""ctypes binding of libpgbsc.so (the C ABI declared in include/pgbsc.h).

The library is the product: there is no Python fallback.  Loading fails
loudly when the shared object is missing, and every call that needs the GPU
raises when no HIP device is usable.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libpgbsc.so")

PG_OK = 0
PG_FP32 = 0
PG_FP64 = 1
PG_FLAG_FIX_2VERTEX = 1
PG_FLAG_UNPRUNED = 2
ERRORS = {-1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "bad state",
          -5: "I/O error", -6: "out of range"}

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


ESTIMATE_FIELDS = ("peak_mem_gb", "pgbsc_spmv_gflop", "pgbsc_fma_gflop", "pgbsc_spmv_gbytes", "pgbsc_fma_gbytes",
                   "fascia_gflop", "fascia_gbytes", "pruned_fascia_gflop", "pruned_fascia_gbytes",
                   "workload_n", "workload_nnz", "memload_n", "memload_nnz")


class PgEstimate(C.Structure):
    """struct pg_estimate (pgbsc.h): thirteen doubles in declaration order."""
    _fields_ = [(f, C.c_double) for f in ESTIMATE_FIELDS]


class PGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} [{ERRORS.get(code, code)}]")
        self.code = code


def _share_hip_runtime() -> None:
    """Load PyTorch before libpgbsc.so when PyTorch is installed.

    The PyTorch-ROCm wheel bundles its own libamdhip64 / libhsa-runtime64,
    which its libraries request under unversioned names.  If libpgbsc.so came
    first, the system runtime (soname libamdhip64.so.7) would be loaded and
    PyTorch would then load its bundled copy beside it: two HSA runtimes in
    one process, and the second one finds no GPU.  Loading PyTorch first makes
    libpgbsc.so's libamdhip64.so.7 resolve to the copy already in the process
    (same soname), so the library and PyTorch share one runtime and device
    pointers pass freely between them.  Set PGBSC_SYSTEM_HIP=1 to skip this
    (processes that never touch PyTorch's GPU side)."""
    import importlib.util
    if os.environ.get("PGBSC_SYSTEM_HIP") == "1" or importlib.util.find_spec("torch") is None:
        return
    import torch  # noqa: F401  (loads libtorch_hip and PyTorch's HIP runtime; no GPU init)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libpgbsc.so not built at {LIB_PATH}: run __graft_entry__.build() "
                          f"or make -C subgraph_amd/csrc")
    _share_hip_runtime()
    L = C.CDLL(LIB_PATH)
    sig = {
        "pg_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
        "pg_destroy": (None, [_vp]),
        "pg_last_error": (C.c_char_p, [_vp]),
        "pg_set_option": (C.c_int, [_vp, C.c_char_p, C.c_int64]),
        "pg_get_option": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int64)]),
        "pg_graph_upload": (C.c_int, [_vp, C.c_int32, C.c_int64, _i32p, _i32p, _vp]),
        "pg_graph_set_weights": (C.c_int, [_vp, _vp, C.c_int]),
        "pg_graph_stats": (C.c_int, [_vp] + [C.POINTER(C.c_int64)] * 4),
        "pg_plan_create": (C.c_int, [_vp, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                     C.POINTER(_vp)]),
        "pg_plan_from_edges": (C.c_int, [_vp, C.c_int, C.c_int, _i32p, _i32p, C.POINTER(_vp)]),
        "pg_plan_destroy": (None, [_vp]),
        "pg_plan_info": (C.c_int, [_vp] + [C.POINTER(C.c_int)] * 4),
        "pg_colors_fill": (None, [C.c_uint64, C.c_int64, C.c_int, C.c_int32, C.c_int, _u8p]),
        "pg_count": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_uint64, C.c_int, C.c_uint32, _f64p]),
        "pg_count_multi": (C.c_int, [_vp, C.c_int, _vp, _i32p, _vp, C.c_int, C.c_uint32, _f64p]),
        "pg_count_grad": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_uint64, C.c_int, C.c_uint32, _vp, _f64p,
                                    _vp, C.c_int]),
        "pg_last_timing": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "pg_last_timing_kernel": (C.c_int, [_vp, C.c_char_p, C.c_int]),
        "pg_last_timing_entry": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                           C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "pg_count_noalloc": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_uint64, C.c_int, C.c_uint32,
                                       C.POINTER(C.c_double)]),
        "pg_last_timing_sub": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "pg_rmat_edges": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.c_int64, _vp, _vp, C.POINTER(C.c_int64)]),
        "pg_csc_from_edges": (C.c_int, [C.c_int32, C.c_int64, _i32p, _i32p, _i32p, _vp]),
        "pg_edges_remove_holes": (C.c_int64, [C.c_int64, C.c_int64, _i32p, _i32p]),
        "pg_template_automorphisms": (C.c_int64, [C.c_int, C.c_int, _i32p, _i32p]),
        "pg_template_sizes": (C.c_int, [C.c_int, C.c_int, _i32p, _i32p, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "pg_template_tables": (C.c_int, [C.c_int, C.c_int, _i32p, _i32p] + [_i32p] * 6),
        "pg_estimate_model": (C.c_int, [C.c_int64, C.c_int64, C.c_int, C.c_int, _i32p, _i32p, _i32p,
                                        C.POINTER(PgEstimate)]),
        "pg_plan_footprint": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list[str]:
    """Names the C ABI header declares (checked against the shared object)."""
    return ["pg_create", "pg_destroy", "pg_last_error", "pg_set_option", "pg_get_option",
            "pg_graph_upload", "pg_graph_set_weights", "pg_graph_stats", "pg_plan_create",
            "pg_plan_from_edges", "pg_plan_destroy", "pg_plan_info", "pg_colors_fill", "pg_count", "pg_count_noalloc",
            "pg_count_multi", "pg_count_grad", "pg_last_timing", "pg_last_timing_kernel",
            "pg_last_timing_entry", "pg_last_timing_sub", "pg_rmat_edges", "pg_csc_from_edges", "pg_edges_remove_holes",
            "pg_template_sizes", "pg_template_tables", "pg_estimate_model", "pg_plan_footprint",
            "pg_template_automorphisms"]
