"""ctypes binding of libuasd_hip.so (declared in include/uasd_hip.h).

The library is built in-tree by ``uppasd_amd.build.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuasd_hip.so")

UASD_OK = 0
UASD_FP64, UASD_FP32 = 0, 1
UASD_NOISE_NONE, UASD_NOISE_PHILOX, UASD_NOISE_INJECTED = 0, 1, 2
UASD_SDE_MIDPOINT, UASD_SDE_HEUN, UASD_SDE_DEPONDT = 1, 4, 5
UASD_SDE_SIB, UASD_SDE_SPHMID = 6, 7

ERRORS = {-1: "argument", -2: "hip", -3: "out of memory", -4: "unsupported", -5: "state",
          -6: "no device"}

# Every extern "C" entry point of include/uasd_hip.h.
EXPORTS = (
    "uasd_last_error", "uasd_version", "uasd_device_count", "uasd_ctx_create",
    "uasd_ctx_destroy", "uasd_ctx_set_stream", "uasd_ctx_stream", "uasd_ctx_sync",
    "uasd_set_exchange", "uasd_set_dm", "uasd_set_anisotropy", "uasd_set_external_field",
    "uasd_set_llg", "uasd_set_noise", "uasd_set_moments", "uasd_get_moments",
    "uasd_effective_field_host", "uasd_evolve_first_host", "uasd_evolve_second_host",
    "uasd_sd_steps", "uasd_effective_field", "uasd_mc_sweeps", "uasd_measure",
    "uasd_get_coloring", "uasd_set_kernel_timing", "uasd_get_kernel_timing",
    "uasd_set_tuning", "uasd_get_tuning", "uasd_autotune_llg", "uasd_autotune_mc",
    "uasd_set_temprescale", "uasd_set_segments", "uasd_measure_segments", "uasd_energy_terms",
    "uasd_energy_terms_noalloc",
    "uasd_accumulate", "uasd_get_accumulators", "uasd_reset_accumulators", "uasd_set_graph_steps",
    "uasd_index_bytes", "uasd_llg_tile", "uasd_set_site_order", "uasd_set_lattice_dims",
    "uasd_has_stencil",
    "uasd_build_pair_list", "uasd_set_tune_cache",
)


class UasdError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)} error: {msg}")
        self.code = code


_lib = None
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        L.uasd_last_error.restype = C.c_char_p
        L.uasd_ctx_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int]
        L.uasd_ctx_destroy.argtypes = [vp]
        L.uasd_ctx_set_stream.argtypes = [vp, vp]
        L.uasd_ctx_stream.argtypes = [vp]
        L.uasd_ctx_stream.restype = vp
        L.uasd_ctx_sync.argtypes = [vp]
        L.uasd_set_exchange.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.uasd_set_dm.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.uasd_set_anisotropy.argtypes = [vp, vp, vp, vp, vp]
        L.uasd_set_external_field.argtypes = [vp, vp]
        L.uasd_set_llg.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_double]
        L.uasd_set_noise.argtypes = [vp, C.c_int, C.c_uint64, vp]
        L.uasd_set_moments.argtypes = [vp, vp, vp]
        L.uasd_get_moments.argtypes = [vp, vp, vp]
        L.uasd_effective_field_host.argtypes = [vp, vp, vp, vp, vp, dp]
        L.uasd_evolve_first_host.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.uasd_evolve_second_host.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.uasd_sd_steps.argtypes = [vp, C.c_int, C.c_int64]
        L.uasd_effective_field.argtypes = [vp, vp, dp]
        L.uasd_mc_sweeps.argtypes = [vp, C.c_char, vp, C.c_int64, C.c_uint64]
        L.uasd_measure.argtypes = [vp, vp, vp]
        L.uasd_get_coloring.argtypes = [vp, ip, vp]
        L.uasd_set_kernel_timing.argtypes = [vp, C.c_int]
        L.uasd_get_kernel_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), dp]
        L.uasd_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.uasd_set_temprescale.argtypes = [vp, C.c_double]
        L.uasd_set_segments.argtypes = [vp, C.c_int, vp]
        L.uasd_measure_segments.argtypes = [vp, vp, vp]
        L.uasd_accumulate.argtypes = [vp, C.c_int]
        L.uasd_energy_terms.argtypes = [vp, vp]
        L.uasd_energy_terms_noalloc.argtypes = [vp, vp]
        L.uasd_get_accumulators.argtypes = [vp, vp, C.POINTER(C.c_int64)]
        L.uasd_reset_accumulators.argtypes = [vp]
        L.uasd_set_graph_steps.argtypes = [vp, C.c_int]
        L.uasd_index_bytes.argtypes = [vp, ip, ip]
        L.uasd_llg_tile.argtypes = [vp, ip]
        L.uasd_set_site_order.argtypes = [vp, C.c_int]
        L.uasd_set_lattice_dims.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.uasd_has_stencil.argtypes = [vp, ip, vp]
        L.uasd_get_tuning.argtypes = [vp, C.c_int, ip, ip]
        L.uasd_autotune_llg.argtypes = [vp, C.c_int, C.c_int, dp]
        L.uasd_autotune_mc.argtypes = [vp, C.c_char, vp, C.c_int, dp]
        L.uasd_set_tune_cache.argtypes = [C.c_int, C.c_char_p]
        L.uasd_build_pair_list.argtypes = [C.c_int] * 5 + [vp] * 6 + [C.c_int, C.c_int] + \
            [vp, vp, vp] + [C.c_int, C.c_int] + [vp, vp, vp] + [ip, vp, vp, vp]
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != UASD_OK:
        msg = lib().uasd_last_error()
        raise UasdError(code, where, msg.decode() if msg else "")


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the engine must be C-contiguous"
    return a.ctypes.data


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)
