"""The C-ABI library: it loads, exports every symbol include/sstmc.h declares, and validates
configs on the host (no GPU needed for these)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from temporalspecklescode_amd import _native, config, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sstmc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sstmc_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(native):
    syms = declared_symbols()
    assert len(syms) >= 19
    missing = [s for s in syms if not hasattr(native, s)]
    assert not missing, missing
    assert set(syms) == set(_native.EXPORTS)


def test_error_strings(native):
    for code in range(11):
        assert native.sstmc_error_string(code)
    assert native.sstmc_error_string(7) == b"Invalid config"


def _create(cfg, dims):
    cc = config.CConfig(cfg, dims)
    h = C.c_void_p()
    rc = _native.lib().sstmc_create(cc.ptr(), C.byref(h))
    if rc == 0:
        _native.lib().sstmc_destroy(h)
    return rc


def test_invalid_configs_rejected_before_device_use(native):
    base = W.c1(paths=64, batch=64)
    # material 0 inside the grid (MatlabInterface.cu:975-992): raised by the python mirror
    bad = dict(base); bad["medium"] = dict(base["medium"]); bad["medium"]["materialGrid"] = np.zeros((1, 1), np.uint32)
    with pytest.raises(ValueError):
        config.CConfig(bad, 2)
    # the C ABI checks the same without the python mirror: patch the struct directly
    cc = config.CConfig(base, 2)
    zero = np.zeros(1, np.uint32)
    cc.struct.medium.materialGrid = zero.ctypes.data_as(C.POINTER(C.c_uint32))
    h = C.c_void_p()
    assert native.sstmc_create(cc.ptr(), C.byref(h)) == 7
    # negative D (MatlabInterface.cu:1293-1297)
    cc = config.CConfig(base, 2)
    negd = np.array([0.0, -1.0])
    cc.struct.sampler.D = negd.ctypes.data_as(C.POINTER(C.c_double))
    assert native.sstmc_create(cc.ptr(), C.byref(h)) == 7
    # coherent backscattering with the correlation output is not built: NOT_SUPPORTED
    cb = dict(base); cb["tracer"] = {"type": "nee", "isCBS": True, "output": "correlation"}
    assert _create(cb, 2) == 5
    cb["tracer"] = {"type": "nee", "isCBS": True}
    assert native.sstmc_validate(config.CConfig(cb, 2).ptr()) == 0
    # non-increasing axis
    cc = config.CConfig(base, 2)
    ax = np.array([1.0, 0.0])
    cc.struct.medium.xAxis = ax.ctypes.data_as(C.POINTER(C.c_double))
    assert native.sstmc_create(cc.ptr(), C.byref(h)) == 7
    # wrong dims
    cc = config.CConfig(base, 2)
    cc.struct.dims = 4
    assert native.sstmc_create(cc.ptr(), C.byref(h)) == 7


def test_python_mirror_messages():
    base = W.c1(paths=64, batch=64)
    bad = dict(base); bad["sampler"] = dict(base["sampler"]); bad["sampler"]["D"] = [0.0, -1.0]
    with pytest.raises(ValueError, match="non-negative"):
        config.CConfig(bad, 2)
    bad = dict(base); bad["scattering"] = {"type": ["Mie", "Mie"]}
    with pytest.raises(ValueError, match="Unsupported scattering"):
        config.CConfig(bad, 2)
    bad = dict(base); bad["scattering"] = {"type": ["Tabular", "Tabular"], "amplitudeFunction": [[1.0], [1.0]]}
    with pytest.raises(ValueError, match="Tabular amplitude function of material 0"):
        config.CConfig(bad, 2)
