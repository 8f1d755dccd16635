"""Continuous-batching request loop (C ABI ``sstmc_server_*``).

Requests of different sizes, media and precisions are admitted by HBM footprint, their batches
are interleaved on the device, and results stay on the device until fetched. Each result is
bit-identical to the same request traced alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .config import CConfig, sizes


class Server:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_inflight: int = 4):
        h = C.c_void_p()
        _native.check(_native.lib().sstmc_server_create(device, hbm_budget_bytes, max_inflight, C.byref(h)),
                      "sstmc_server_create")
        self._h = h
        self._meta = {}

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().sstmc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, config: dict, dims: int) -> int:
        cc = CConfig(config, dims)
        rid = C.c_uint64()
        _native.check(_native.lib().sstmc_submit(self._h, cc.ptr(), C.byref(rid)), "sstmc_submit")
        nl, nv, nt, _ = sizes(config, dims)
        self._meta[rid.value] = (nl, nv, nt)
        return rid.value

    def poll(self, rid: int) -> bool:
        rc = _native.lib().sstmc_poll(self._h, C.c_uint64(rid))
        if rc == 10:
            return False
        _native.check(rc, "sstmc_poll")
        return True

    def wait(self, rid: int):
        _native.check(_native.lib().sstmc_wait(self._h, C.c_uint64(rid)), "sstmc_wait")

    def fetch(self, rid: int):
        nl, nv, nt = self._meta[rid]
        u = np.zeros(2 * nl * nv * nt)
        it = C.c_double()
        _native.check(_native.lib().sstmc_fetch(self._h, C.c_uint64(rid), u.ctypes.data_as(C.c_void_p), C.byref(it)),
                      "sstmc_fetch")
        return (u[0::2] + 1j * u[1::2]).reshape(nt, nv, nl).transpose(2, 1, 0), it.value

    def stats(self) -> dict:
        s = _native.ServerStats()
        _native.check(_native.lib().sstmc_server_stats_get(self._h, C.byref(s)), "sstmc_server_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}
