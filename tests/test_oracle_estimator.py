"""Estimator-level checks of the FP64 CPU restatement: the reference's own known answer (the
MCX temporal-correlation curve of examples/mcxValidation3D.m) and invariants of the estimator."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from temporalspecklescode_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mcx_validation.json")


def mcx_cfg(batch, seed):
    r = json.load(open(GOLD))
    return r, W.mcx_validation(r["t"], r["D"], r["MFP"], r["R"], r["h"], r["dims"], r["viSep"], batch=batch, seed=seed)


def test_mcx_temporal_correlation_curve():
    """examples/mcxValidation3D.m overlays abs(u/iter) of the correlation estimator on mcx_res(t);
    here the FP64 restatement's curve must agree within 4 standard errors at every t."""
    res = []
    for s in range(8):
        r, cfg = mcx_cfg(20000, 200 + s)
        u, it, _ = O.render(cfg, 3)
        res.append(u[0, 0, :].real / it)
    res = np.array(res)
    mcx = np.array(r["mcx_res"])
    m = res.mean(0)
    se = res.std(0, ddof=1) / np.sqrt(len(res))
    z = (m - mcx) / se
    assert np.all(np.abs(z) < 4), z
    # the curve decays: C(0) > C(t_max)
    assert m[0] > 1.8 * m[-1]


def test_field_mean_is_zero_and_correlation_at_zero_lag_is_intensity():
    """E[u] = 0 (uniform random phase per vertex), and the per-sub-path correlation estimator at
    lag 0 equals the mean field intensity E|u|^2 over renderings."""
    cfg = W.c1(paths=512, batch=512)
    cfg["illumination"]["direction"] = {"x": [0.0, 0.1], "y": [1.0, 1.0]}
    cfg["view"]["direction"] = {"x": [0.0, 0.2], "y": [1.0, 1.0]}
    us, its = [], []
    for s in range(60):
        cfg["simulation"]["seed"] = 1000 + s
        u, it, _ = O.render(cfg, 2)
        us.append(u[:, :, 0])
        its.append(it)
    us = np.array(us)
    inten = np.mean(np.abs(us) ** 2, axis=0)
    mean = np.mean(us, axis=0)
    assert np.all(np.abs(mean) < 4 * np.sqrt(inten / len(us)))
    c = dict(cfg)
    c["tracer"] = {"type": "nee", "output": "correlation"}
    cs = []
    for s in range(60):
        c["simulation"] = dict(cfg["simulation"], seed=5000 + s)
        u, it, _ = O.render(c, 2)
        cs.append(u[:, :, 0].real)
    cs = np.array(cs)
    cm, cse = cs.mean(0), cs.std(0, ddof=1) / np.sqrt(len(cs))
    ise = np.std(np.abs(us) ** 2, axis=0, ddof=1) / np.sqrt(len(us))
    assert np.all(np.abs(cm - inten) < 4 * np.sqrt(cse ** 2 + ise ** 2))


def test_point_source_reciprocity():
    """Swapping point illuminations and point sensors leaves the mean intensity unchanged
    (reciprocity of the transport with the reference's symmetric g for point sources). The
    sources sit outside the box: the reference's NEE walk integrates to the box boundary rather
    than to the point (HeterogeneousMedium.cuh:453-550), which is the physical attenuation only
    for points outside the medium."""
    n = 16
    edges = np.arange(-4, 5, dtype=float)
    grid = W.random_materials((8, 8, 8), 9, 3)
    def cfg(src, dst, seed):
        return {"simulation": {"batchSize": 2000, "seed": seed, "lambda": 0.5},
                "medium": {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "zAxis": edges,
                           "materialGrid": grid, "sigs": [0.0, 0.3, 0.6, 0.9], "siga": [0.0, 0.01, 0.02, 0.03]},
                "scattering": {"type": ["Isotropic"] * 4},
                "illumination": {"type": "PointSource", "location": src},
                "view": {"type": "PointSource", "location": dst},
                "sampler": {"type": "TemporalSampler", "t": [0.0], "D": np.zeros(4),
                            "U": {"x": np.zeros(4), "y": np.zeros(4), "z": np.zeros(4)}},
                "tracer": {"type": "nee", "output": "correlation"}}
    a = {"x": [-1.0], "y": [0.5], "z": [-4.5]}
    b = {"x": [1.5], "y": [-0.5], "z": [4.5]}
    fw, bw = [], []
    for s in range(n):
        u, it, _ = O.render(cfg(a, b, 10 + s), 3)
        fw.append(u[0, 0, 0].real / it)
        u, it, _ = O.render(cfg(b, a, 50 + s), 3)
        bw.append(u[0, 0, 0].real / it)
    fw, bw = np.array(fw), np.array(bw)
    se = np.sqrt(fw.var(ddof=1) / n + bw.var(ddof=1) / n)
    assert abs(fw.mean() - bw.mean()) < 4 * se, (fw.mean(), bw.mean(), se)
