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


def test_correlation_sampler_config_checks(native):
    """TemporalCorrelationSampler: illumination2 / view2 must be present and as large as
    illumination / view (MatlabInterface.cu:1313-1331); u1 and u2 share one source type here; with
    isCBS it takes the correlation pairings CBS01-CBS04."""
    base = W.c1(paths=64, batch=64)
    cs = dict(base, sampler=dict(base["sampler"], type="TemporalCorrelationSampler"),
              illumination2=base["illumination"], view2=base["view"])
    assert native.sstmc_validate(config.CConfig(cs, 2).ptr()) == 0
    cc = config.CConfig(cs, 2)
    cc.struct.view2.type = 0                          # absent in the C struct
    assert native.sstmc_validate(cc.ptr()) == 7
    cc = config.CConfig(cs, 2)
    cc.struct.illumination2.num = 3                   # sizes differ
    assert native.sstmc_validate(cc.ptr()) == 7
    pt = dict(cs, view2={"type": "PointSource", "location": {"x": np.zeros(16), "y": np.full(16, 30.0)}})
    assert native.sstmc_validate(config.CConfig(pt, 2).ptr()) == 5   # u2 of another type
    cb = dict(cs, tracer={"type": "nee", "isCBS": True})
    assert native.sstmc_validate(config.CConfig(cb, 2).ptr()) == 0
    with pytest.raises(ValueError, match="illuminationsHandle2"):
        config.CConfig({k: v for k, v in cs.items() if k != "view2"}, 2)
    with pytest.raises(ValueError, match="Available sampler types are TemporalSampler and TemporalCorrelationSampler"):
        config.CConfig(dict(base, sampler=dict(base["sampler"], type="Other")), 2)


def test_gaussian_beam_config_checks(native):
    """GaussianBeam (MatlabInterface.cu:1153-1190): 2D or 3D, a direction per focal point, aperture > 0;
    built for GaussianBeam illumination and view together, vMF mixtures of at most 16 lobes, the
    TemporalSampler without coherent backscattering."""
    cfg = W.confocal(focal_x=[0.0, 2.0], focal_z=[8.0], tau_x=[0.0], batch=64, full_iteration=1)
    assert native.sstmc_validate(config.CConfig(cfg, 3).ptr()) == 0
    cc = config.CConfig(cfg, 3)
    cc.struct.illumination.aperture = 0.0
    assert native.sstmc_validate(cc.ptr()) == 7
    ff = dict(cfg, view={"type": "FarField", "direction": {"x": [0.0], "y": [0.0], "z": [1.0]}})
    assert native.sstmc_validate(config.CConfig(ff, 3).ptr()) == 5
    corr = dict(cfg, tracer={"type": "nee", "output": "correlation"})
    assert native.sstmc_validate(config.CConfig(corr, 3).ptr()) == 0
    cbs = dict(cfg, tracer={"type": "nee", "isCBS": True})
    assert native.sstmc_validate(config.CConfig(cbs, 3).ptr()) == 5
    hg = dict(cfg, scattering={"type": ["HG"] * 5, "amplitudeFunction": [0.8] * 5})
    assert native.sstmc_validate(config.CConfig(hg, 3).ptr()) == 5
    flat = dict(cfg, medium=dict(cfg["medium"], materialGrid=np.array([[1, 2]], np.uint32)))
    assert native.sstmc_validate(config.CConfig(flat, 2).ptr()) == 0   # 2D beams (x, y only)


def test_config_pointer_keeps_its_arrays(native):
    """A pointer from a temporary CConfig holds the config (and the arrays its struct points into)
    until the pointer is released: validation through it after garbage collection and new
    allocations still sees the original arrays."""
    import gc
    p = config.CConfig(W.gauss_transmission_matrix(pixels=9, dp=1.0, batch=16, full_iteration=1), 3).ptr()
    gc.collect()
    junk = [np.full(1 << 16, 7.0) for _ in range(64)]
    assert native.sstmc_validate(p) == 0
    assert p.contents.illumination.num == 81 and p.contents.illumination.aperture == 0.4
    del junk
