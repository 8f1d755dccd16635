"""MI355X-native next-event-estimation speckle tracer (sstmc2D / sstmc3D).

``sstmc2D(config)`` / ``sstmc3D(config)`` mirror the reference's MEX entry
``[u, iter] = sstmc3D(config)`` (MatlabInterface.cu:1565-1825): ``config`` is a dict with the
same fields as the MATLAB struct, ``u`` is complex with shape (Nl, Nv, Nt, renderingsNum) and
``iter`` has shape (renderingsNum,). Normalisation (u / sqrt(iter) for fields, C / iter for
correlations) stays with the caller, as in the reference (README.md:365-373).

The work runs in libsstmc.so (hand-written HIP for gfx950) behind the C ABI of include/sstmc.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .config import CConfig, sizes

__all__ = ["sstmc2D", "sstmc3D", "Tracer", "render"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def render(config: dict, dims: int):
    """``[u, iter] = sstmc{dims}D(config)``."""
    cc = CConfig(config, dims)
    nl, nv, nt, rn = sizes(config, dims)
    u = np.zeros(2 * nl * nv * nt * rn)
    it = np.zeros(rn)
    _native.check(_native.lib().sstmc_render(cc.ptr(), _p(u), _p(it)), "sstmc_render")
    uc = (u[0::2] + 1j * u[1::2]).reshape(rn, nt, nv, nl).transpose(3, 2, 1, 0)
    return uc, it


def sstmc2D(config: dict):
    return render(config, 2)


def sstmc3D(config: dict):
    return render(config, 3)


class Tracer:
    """One configured tracer (Simulation + Medium + Sources + Sampler + NEE tracer)."""

    def __init__(self, config: dict, dims: int):
        self._cc = CConfig(config, dims)
        self.dims = dims
        self.L, self.V, self.T, _ = sizes(config, dims)
        h = C.c_void_p()
        _native.check(_native.lib().sstmc_create(self._cc.ptr(), C.byref(h)), "sstmc_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().sstmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def trace(self, seed: int | None = None):
        """One rendering. Returns (u[l, v, t] complex128, iter)."""
        L = _native.lib()
        if seed is not None:
            _native.check(L.sstmc_update_seed(self._h, C.c_uint64(seed)), "sstmc_update_seed")
        u = np.zeros(2 * self.L * self.V * self.T)
        it = C.c_double()
        _native.check(L.sstmc_trace(self._h, _p(u), C.byref(it)), "sstmc_trace")
        return (u[0::2] + 1j * u[1::2]).reshape(self.T, self.V, self.L).transpose(2, 1, 0), it.value

    def timing(self) -> dict:
        t = _native.Timing()
        _native.check(_native.lib().sstmc_last_timing(self._h, C.byref(t)), "sstmc_last_timing")
        return {k: getattr(t, k) for k, _ in t._fields_}

    @property
    def zero_contribution_paths(self) -> int:
        return int(_native.lib().sstmc_zero_contribution_paths(self._h))

    @property
    def photons_budget(self) -> int:
        return int(_native.lib().sstmc_photons_budget(self._h))

    def debug_batch(self, begin: int, n: int) -> dict:
        """Vertices and per-path terms of paths [begin, begin+n), plus the batch's u."""
        L = _native.lib()
        nv = C.c_uint64()
        nul = [None] * 14
        _native.check(L.sstmc_debug_batch(self._h, C.c_uint64(begin), C.c_uint32(n), C.byref(nv), *nul[:13]),
                      "sstmc_debug_batch")
        m = nv.value
        T, Lh, Vv = self.T, self.L, self.V
        out = dict(k=np.zeros(n, np.uint32), misses=np.zeros(n, np.uint32), off=np.zeros(n + 1, np.uint64),
                   pos=np.zeros((m, 3), np.float32), mat=np.zeros(m, np.uint32), rho=np.zeros(m, np.float32),
                   dxt=np.zeros((m, T, 3), np.float32), len=np.zeros((m, T)), a=np.zeros((n, T, Lh, 2)),
                   b=np.zeros((n, T, Vv, 2)), sa=np.zeros((n, T, Lh, 2)), sb=np.zeros((n, T, Vv, 2)),
                   u=np.zeros(2 * Lh * Vv * T))
        _native.check(L.sstmc_debug_batch(self._h, C.c_uint64(begin), C.c_uint32(n), C.byref(nv),
                                          _p(out["k"]), _p(out["misses"]), _p(out["off"]), _p(out["pos"]),
                                          _p(out["mat"]), _p(out["rho"]), _p(out["dxt"]), _p(out["len"]),
                                          _p(out["a"]), _p(out["b"]), _p(out["sa"]), _p(out["sb"]), _p(out["u"])),
                      "sstmc_debug_batch")
        for key in ("a", "b", "sa", "sb"):
            out[key] = out[key][..., 0] + 1j * out[key][..., 1]
        u = out["u"]
        out["u"] = (u[0::2] + 1j * u[1::2]).reshape(T, Vv, Lh).transpose(2, 1, 0)
        # frame positions in FP64: central FP32 position + FP32 displacement, summed exactly
        out["xt"] = out["pos"].astype(np.float64)[:, None, :] + out["dxt"].astype(np.float64)
        return out

    def debug_cbs(self, n: int):
        """Coherent-backscattering terms of the last debug_batch(begin, n): Ar [n, Nt, Nv] (view
        connected at x1) and Br [n, Nt, Nl] (illumination connected at the later vertices)."""
        ar = np.zeros((n, self.T, self.V, 2))
        br = np.zeros((n, self.T, self.L, 2))
        _native.check(_native.lib().sstmc_debug_cbs(self._h, C.c_uint32(n), _p(ar), _p(br)), "sstmc_debug_cbs")
        return ar[..., 0] + 1j * ar[..., 1], br[..., 0] + 1j * br[..., 1]

    def debug_tau_ref(self) -> float:
        """Reference optical depth of the relative attenuation cut in the last batch traced."""
        r = C.c_float()
        _native.check(_native.lib().sstmc_debug_tau_ref(self._h, C.byref(r)), "sstmc_debug_tau_ref")
        return float(r.value)

    def debug_contract(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """u[l, v, t] = sum_p a[p, t, l] b[p, t, v] with this tracer's contraction kernel."""
        n = a.shape[0]
        at = np.ascontiguousarray(np.transpose(a, (1, 0, 2)))  # [t][p][l]
        bt = np.ascontiguousarray(np.transpose(b, (1, 0, 2)))
        ai = np.ascontiguousarray(np.stack([at.real, at.imag], axis=-1), np.float64)
        bi = np.ascontiguousarray(np.stack([bt.real, bt.imag], axis=-1), np.float64)
        u = np.zeros(2 * self.L * self.V * self.T)
        _native.check(_native.lib().sstmc_debug_contract(self._h, C.c_uint32(n), _p(ai), _p(bi), _p(u)),
                      "sstmc_debug_contract")
        return (u[0::2] + 1j * u[1::2]).reshape(self.T, self.V, self.L).transpose(2, 1, 0)
