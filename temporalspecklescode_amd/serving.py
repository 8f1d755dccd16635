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

    def preload(self, config: dict, dims: int):
        """Builds the medium of `config` on the device and keeps it for later requests."""
        _native.check(_native.lib().sstmc_server_preload(self._h, CConfig(config, dims).ptr()), "sstmc_server_preload")

    def submit(self, config: dict, dims: int) -> int:
        cc = CConfig(config, dims)
        rid = C.c_uint64()
        _native.check(_native.lib().sstmc_submit(self._h, cc.ptr(), C.byref(rid)), "sstmc_submit")
        self._meta[rid.value] = sizes(config, dims)
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
        """(u, iter) like ``sstmc2D/3D``: u of shape (Nl, Nv, Nt) and iter a float for one
        rendering; (Nl, Nv, Nt, renderingsNum) and an array of renderingsNum values otherwise."""
        nl, nv, nt, rn = self._meta[rid]
        u = np.zeros(2 * nl * nv * nt * rn)
        it = np.zeros(rn)
        _native.check(_native.lib().sstmc_fetch(self._h, C.c_uint64(rid), u.ctypes.data_as(C.c_void_p),
                                                it.ctypes.data_as(C.c_void_p)), "sstmc_fetch")
        u = (u[0::2] + 1j * u[1::2]).reshape(rn, nt, nv, nl).transpose(3, 2, 1, 0)
        if rn == 1:
            return u[..., 0], float(it[0])
        return u, it

    def stats(self) -> dict:
        s = _native.ServerStats()
        _native.check(_native.lib().sstmc_server_stats_get(self._h, C.byref(s)), "sstmc_server_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}
