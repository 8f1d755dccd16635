"""ctypes binding of libsstmc.so (the C ABI in include/sstmc.h).

The HIP library is the only compute path: if it is missing this module raises, it never falls
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsstmc.so")

EXPORTS = [
    "sstmc_create", "sstmc_trace", "sstmc_update_seed", "sstmc_zero_contribution_paths",
    "sstmc_photons_budget", "sstmc_sizes", "sstmc_destroy", "sstmc_error_string", "sstmc_render",
    "sstmc_last_timing", "sstmc_debug_batch", "sstmc_debug_contract", "sstmc_server_create",
    "sstmc_submit", "sstmc_poll", "sstmc_wait", "sstmc_fetch", "sstmc_server_stats_get",
    "sstmc_server_destroy", "sstmc_validate", "sstmc_debug_cbs", "sstmc_debug_tau_ref",
    "sstmc_contract_noalloc", "sstmc_server_preload",
]


class Timing(C.Structure):
    _fields_ = [("transport_ms", C.c_double), ("nee_ms", C.c_double), ("contract_ms", C.c_double),
                ("single_ms", C.c_double), ("total_ms", C.c_double), ("vertices", C.c_uint64),
                ("dda_steps", C.c_uint64), ("batches", C.c_uint32), ("nee_view_ms", C.c_double),
                ("dda_steps_view", C.c_uint64), ("graph_replays", C.c_uint32),
                ("pipeline_sets", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("paths_per_s", C.c_double), ("contributions_per_s", C.c_double),
                ("hbm_in_use", C.c_uint64), ("hbm_budget", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    """Loads libsstmc.so (build it with ``python -m temporalspecklescode_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsstmc.so is not built (%s); run __graft_entry__.build() or "
                           "python temporalspecklescode_amd/build.py" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sstmc_create.argtypes = [vp, C.POINTER(vp)]
    L.sstmc_trace.argtypes = [vp, vp, vp]
    L.sstmc_update_seed.argtypes = [vp, C.c_uint64]
    L.sstmc_zero_contribution_paths.argtypes = [vp]
    L.sstmc_zero_contribution_paths.restype = C.c_uint64
    L.sstmc_photons_budget.argtypes = [vp]
    L.sstmc_photons_budget.restype = C.c_uint64
    L.sstmc_sizes.argtypes = [vp, vp, vp, vp]
    L.sstmc_destroy.argtypes = [vp]
    L.sstmc_destroy.restype = None
    L.sstmc_error_string.argtypes = [C.c_int]
    L.sstmc_error_string.restype = C.c_char_p
    L.sstmc_render.argtypes = [vp, vp, vp]
    L.sstmc_last_timing.argtypes = [vp, vp]
    L.sstmc_debug_batch.argtypes = [vp, C.c_uint64, C.c_uint32] + [vp] * 14
    L.sstmc_debug_contract.argtypes = [vp, C.c_uint32, vp, vp, vp]
    L.sstmc_debug_tau_ref.argtypes = [vp, vp]
    L.sstmc_contract_noalloc.argtypes = [C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, vp]
    L.sstmc_server_create.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(vp)]
    L.sstmc_submit.argtypes = [vp, vp, C.POINTER(C.c_uint64)]
    L.sstmc_server_preload.argtypes = [vp, vp]
    L.sstmc_poll.argtypes = [vp, C.c_uint64]
    L.sstmc_wait.argtypes = [vp, C.c_uint64]
    L.sstmc_fetch.argtypes = [vp, C.c_uint64, vp, vp]
    L.sstmc_server_stats_get.argtypes = [vp, vp]
    L.sstmc_server_destroy.argtypes = [vp]
    L.sstmc_server_destroy.restype = None
    _lib = L
    return L


class SstmcError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().sstmc_error_string(code).decode()
        super().__init__("%s failed (%d): %s" % (where, code, msg))
        self.code = code


def check(code: int, where: str):
    if code != 0:
        raise SstmcError(code, where)
