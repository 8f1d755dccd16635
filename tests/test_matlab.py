"""The MATLAB front end (temporalspecklescode_amd.matlab): a config struct as MATLAB saves it in a
.mat file (cells for scattering.type and scattering.amplitudeFunction, struct cells for vMF
mixtures, a one-material cell, double-valued grids indexed (x, y, z)) loads into a config that
renders exactly as the dict it was written from (FP64 restatement, same seed)."""
import json
import os

import numpy as np
import pytest
import scipy.io as sio

import oracle_lib as O
from temporalspecklescode_amd import matlab as M
from temporalspecklescode_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def _cell(items):
    c = np.empty(len(items), dtype=object)
    for i, x in enumerate(items):
        c[i] = _mat(x)
    return c


def _mat(x):
    """A config value the way MATLAB stores it: cells for lists, structs for dicts, doubles for
    numbers and grids."""
    if isinstance(x, dict):
        return {k: _mat(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        if all(isinstance(e, str) for e in x) or any(isinstance(e, (dict, np.ndarray, list)) for e in x):
            return _cell(list(x))
        return np.asarray(x, np.float64)
    if isinstance(x, np.ndarray) and x.dtype.kind in "ui":
        return x.astype(np.float64)
    return x


def _roundtrip(cfg, tmp_path, cellify=("type", "amplitudeFunction")):
    m = _mat(cfg)
    sc = m["scattering"]
    for k in cellify:                              # the MEX takes both scattering fields as cells
        if k in cfg["scattering"] and not (isinstance(sc[k], np.ndarray) and sc[k].dtype == object):
            v = cfg["scattering"][k]
            sc[k] = _cell(list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v])
    p = os.path.join(str(tmp_path), "cfg.mat")
    sio.savemat(p, {"config": m})
    return M.load_config(p)


def _same_render(a, b, dims):
    ua, ia, _ = O.render(a, dims)
    ub, ib, _ = O.render(b, dims)
    assert ia == ib
    assert np.array_equal(ua, ub)


def test_mcx_config(tmp_path):
    r = json.load(open(os.path.join(HERE, "golden", "mcx_validation.json")))
    cfg = W.mcx_validation(r["t"], r["D"], r["MFP"], r["R"], r["h"], r["dims"], r["viSep"], batch=500, seed=3)
    back = _roundtrip(cfg, tmp_path)
    assert back["scattering"]["type"] == ["HG", "HG", "HG"]
    assert back["scattering"]["amplitudeFunction"] == [0.0, 0.0, 0.0]
    # the reader drops the singleton y axis; the cells keep MATLAB's (column-major) order
    assert np.array_equal(np.ravel(back["medium"]["materialGrid"], order="F"),
                          np.ravel(cfg["medium"]["materialGrid"], order="F"))
    assert M.dims_of(back) == 3
    _same_render(cfg, back, 3)


def test_tabular_2d_and_single_material_cells(tmp_path):
    th = np.linspace(0, 2 * np.pi, 181)
    amp = (1 + 0.5 * np.cos(th)) * np.exp(0.3j * np.sin(th))
    cfg = W.mudiff_cbs(amp, np.linspace(0, 2 * np.pi, 7), [-5.0, 5.0], [-5.0, 5.0], 3.0, batch=300, seed=4)
    back = _roundtrip(cfg, tmp_path)
    assert M.dims_of(back) == 2
    _same_render(cfg, back, 2)
    # one material: {'HG'} and {0.8} come back from the MAT-file as 'HG' and 0.8
    c1 = W.c1(paths=256, batch=256)
    c1["scattering"] = {"type": ["HG"], "amplitudeFunction": [0.8]}
    back = _roundtrip(c1, tmp_path)
    assert back["scattering"] == {"type": ["HG"], "amplitudeFunction": [0.8]}
    _same_render(c1, back, 2)


def test_vmf_struct_cells_and_correlation_sampler(tmp_path):
    cfg = W.c1(paths=256, batch=256)
    mix = {"mixtureMu": [1.0, 4.0], "mixtureC": [-1.0, -2.5], "mixtureAlpha": [0.5, 0.5]}
    cfg["medium"] = dict(cfg["medium"], zAxis=[0.0, 4.0], materialGrid=np.ones((1, 1, 1), np.uint32))
    cfg["illumination"] = {"type": "FarField", "direction": {"x": [0.0], "y": [0.0], "z": [1.0]}}
    cfg["view"] = {"type": "FarField", "direction": {"x": [0.0, 0.1], "y": [0.0, 0.0], "z": [1.0, 1.0]}}
    cfg["scattering"] = {"type": ["vMF", "vMF"], "amplitudeFunction": [mix, mix]}
    cfg["sampler"] = {"type": "TemporalCorrelationSampler", "t": [0.0, 1.0], "D": [0.0, 1e-3],
                      "U": {"x": [0.0, 0.0], "y": [0.0, 0.0], "z": [0.0, 0.0]}}
    cfg["illumination2"] = cfg["illumination"]
    cfg["view2"] = cfg["view"]
    back = _roundtrip(cfg, tmp_path)
    assert isinstance(back["scattering"]["amplitudeFunction"][0], dict)
    _same_render(cfg, back, 3)


def test_check_and_errors(tmp_path, capsys):
    cfg = W.c1(paths=256, batch=256)
    p = os.path.join(str(tmp_path), "c.mat")
    sio.savemat(p, {"config": _mat(dict(cfg, scattering={"type": _cell(["HG", "HG"]),
                                                          "amplitudeFunction": _cell([0.8, 0.8])}))})
    assert M.main([p, "--check"]) == 0
    assert '"valid": true' in capsys.readouterr().out
    assert M.main([p, "--check", "--name", "cfg"]) == 7
    assert "holds no variable 'cfg'" in capsys.readouterr().err
    bad = dict(cfg, scattering={"type": _cell(["HG", "Mie"]), "amplitudeFunction": _cell([0.8, 0.8])})
    sio.savemat(p, {"config": _mat(bad)})
    assert M.main([p, "--check"]) == 7
    assert "Unsupported scattering type" in capsys.readouterr().err
