"""The allocation-free contraction entry point (sstmc_contract_noalloc) against the contraction
entry point it varies (sstmc_debug_contract): same kernels and split plan, so the same bits, on two
calls in a row, the second smaller; and the size limit of its fixed scratch array."""
import ctypes as C

import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer
from temporalspecklescode_amd import _native

pytestmark = pytest.mark.gpu

KIND = {"bf16x3": 0, "fp32": 1, "fp64": 2}


class _Hip:
    """The few HIP runtime calls the test needs for its own device buffers (the entry point under
    test allocates nothing; the test does)."""

    def __init__(self):
        import glob
        import os
        paths = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")]
        paths += sorted(glob.glob("/opt/rocm*/lib/libamdhip64.so*"))
        self.h = next(C.CDLL(q) for q in paths if os.path.exists(q))
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def upload(self, arr):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), arr.nbytes) == 0
        assert self.h.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0   # host to device
        return p

    def zeros(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        assert self.h.hipMemset(p, 0, nbytes) == 0
        return p

    def download(self, p, arr):
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), p, arr.nbytes, 2) == 0   # device to host
        return arr

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def _noalloc(kind, a, b):
    """u[l, v, t] = sum_p a[p, t, l] b[p, t, v] through sstmc_contract_noalloc on device arrays."""
    hip = _Hip()
    P, T, L = a.shape
    V = b.shape[2]
    cdt = np.complex128 if kind == "fp64" else np.complex64
    at = np.ascontiguousarray(np.transpose(a, (1, 0, 2)).astype(cdt))   # [t][p][l]
    bt = np.ascontiguousarray(np.transpose(b, (1, 0, 2)).astype(cdt))
    da, db = hip.upload(at), hip.upload(bt)
    du = hip.zeros(16 * T * V * L)
    rc = _native.lib().sstmc_contract_noalloc(KIND[kind], P, L, V, T, da, db, du, None)
    _native.check(rc, "sstmc_contract_noalloc")
    u = hip.download(du, np.zeros(2 * T * V * L))
    hip.free(da, db, du)
    return (u[0::2] + 1j * u[1::2]).reshape(T, V, L).transpose(2, 1, 0)


@pytest.mark.parametrize("kind", ["bf16x3", "fp32", "fp64"])
def test_noalloc_equals_debug_contract(kind):
    rng = np.random.default_rng(11)
    cdt = np.complex128 if kind == "fp64" else np.complex64
    for (P, L, V, T) in ((4096, 130, 129, 2), (1000, 37, 45, 1)):   # the second call smaller
        print(kind, "P", P, "L", L, "V", V, "T", T, flush=True)
        scale = np.exp(rng.normal(size=(P, 1, 1)))
        a = ((rng.normal(size=(P, T, L)) + 1j * rng.normal(size=(P, T, L))) * scale).astype(cdt)
        b = (rng.normal(size=(P, T, V)) + 1j * rng.normal(size=(P, T, V))).astype(cdt)
        cfg = G.slab_frames(64, nl=L, nv=V, nt=T)
        cfg["precision"] = {"terms": "fp64" if kind == "fp64" else "fp32", "contraction": kind}
        with Tracer(cfg, 3) as tr:
            ref = tr.debug_contract(a.astype(np.complex128), b.astype(np.complex128))
        got = _noalloc(kind, a, b)
        exact = np.stack([a[:, t, :].astype(np.complex128).T @ b[:, t, :].astype(np.complex128)
                          for t in range(T)], axis=2)
        print(kind, "rel err vs FP64 %.3e" % (np.linalg.norm(got - exact) / np.linalg.norm(exact)), flush=True)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_sizes_past_its_scratch():
    lib = _native.lib()
    p = C.c_void_p(1)   # never dereferenced: the size check comes first
    assert lib.sstmc_contract_noalloc(0, 64, 4096, 4096, 1, p, p, p, None) == 11   # 128 MiB per split
    assert lib.sstmc_contract_noalloc(2, 64, 2048, 2049, 1, p, p, p, None) == 11   # complex128: just past 64 MiB
    assert lib.sstmc_contract_noalloc(0, 0, 16, 16, 1, p, p, p, None) == 7
    assert lib.sstmc_contract_noalloc(5, 64, 16, 16, 1, p, p, p, None) == 7
