"""Loader for the native library ``libgrw_fru.so`` (HIP, gfx950).

The library is built in-tree by ``greenstein2002_amd.build`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing, or no GPU is visible when a
context is created, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrw_fru.so")

GRW_FP64 = 0
GRW_MIXED = 1

# Every symbol declared in include/grw_fru.h.
EXPORTS = (
    "grw_last_error",
    "grw_config_default",
    "grw_create",
    "grw_destroy",
    "grw_load_state",
    "grw_get_state",
    "grw_seed",
    "grw_distrib_simfru",
    "grw_calc_flux",
    "grw_calc_avg",
    "grw_save_state",
    "grw_resume_state",
    "grw_simfru_local",
    "grw_calc_fru_flux_local",
    "grw_calc_fru_avg_local",
    "grw_calc_fru_flux_local_noalloc",
    "grw_get_stats",
    "grw_reset_stats",
    "grw_set_timing",
    "grw_set_graphs",
)


class GrwConfig(C.Structure):
    _fields_ = [
        ("nfru", C.c_int32),
        ("precision", C.c_int32),
        ("nfru_scale", C.c_double),
        ("block", C.c_int32),
        ("sequential_sums", C.c_int32),
        ("k23_scale", C.c_double),
        ("gamma_scale", C.c_double),
        ("device", C.c_int32),
        ("lanes_per_fru", C.c_int32),
    ]


class GrwStats(C.Structure):
    _fields_ = [
        ("fru_steps", C.c_uint64),
        ("simfru_calls", C.c_uint64),
        ("launches", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("critical_steps", C.c_uint64),
        ("tail_ms", C.c_double),
        ("timed_launches", C.c_uint64),
        ("slot_steps", C.c_uint64),
    ]


class GrwError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libgrw_fru.so once and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GrwError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    L = C.CDLL(path)
    vp = C.c_void_p
    L.grw_last_error.restype = C.c_char_p
    L.grw_config_default.argtypes = [C.POINTER(GrwConfig)]
    L.grw_config_default.restype = None
    L.grw_create.argtypes = [C.POINTER(GrwConfig), C.POINTER(vp)]
    L.grw_destroy.argtypes = [vp]
    L.grw_load_state.argtypes = [vp] * 7
    L.grw_get_state.argtypes = [vp] * 7
    L.grw_seed.argtypes = [vp, C.c_uint32]
    L.grw_distrib_simfru.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp]
    L.grw_calc_flux.argtypes = [vp, vp, vp]
    L.grw_calc_avg.argtypes = [vp, vp, vp]
    L.grw_save_state.argtypes = [vp]
    L.grw_resume_state.argtypes = [vp]
    L.grw_simfru_local.argtypes = [vp, C.c_double, C.c_double, vp, vp]
    L.grw_calc_fru_flux_local.argtypes = [vp, vp]
    L.grw_calc_fru_avg_local.argtypes = [vp, vp]
    L.grw_calc_fru_flux_local_noalloc.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    L.grw_get_stats.argtypes = [vp, C.POINTER(GrwStats)]
    L.grw_reset_stats.argtypes = [vp]
    L.grw_set_timing.argtypes = [vp, C.c_int]
    L.grw_set_graphs.argtypes = [vp, C.c_int]
    for name in EXPORTS:
        if name not in ("grw_last_error", "grw_config_default"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = load().grw_last_error()
        raise GrwError(f"grw call failed ({rc}): {msg.decode() if msg else ''}")
