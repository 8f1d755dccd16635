"""The TemporalCorrelationSampler on the GPU against the FP64 CPU restatement (TemporalSampler
Correlation.cuh; SourceInterface.cuh:217-352; NextEventEstimation.cuh:1016-1200, 1400-1425): the
u1/u2 mixture sampling of the first vertex, the lag-centred positions of g and f (quirk Q3), the
momentum-transfer decorrelation at x1, x_k and the inner vertices, the scalar-Fs and generic single
scattering, lag chunks, and the reference's MCX known answer (examples/mcxValidation3D.m)."""
import json
import math
import os

import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer, sstmc3D
from temporalspecklescode_amd import workloads as W
from test_gpu_parity import _oracle_u, _rel, _transport_match

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

D4 = [0.0, 0.01, 0.02, 0.005]
U4 = [[0, 0, 0], [0.05, 0.0, 0.02], [-0.03, 0.04, 0.0], [0.0, 0.0, 0.0]]


def _far3(n):
    return G.hetero3_small(n)


def _pt3(n):
    return G.hetero3_small(n, point=True)


def _pt3_iso(n):
    cfg = G.hetero3_small(n, point=True)
    cfg["scattering"] = {"type": ["Isotropic"] * 4}
    return cfg


CASES = {
    # far field, Brownian only, u2 = u1 (the static view-term path, one G per vertex)
    "far3_static_same": (lambda n: G.corr_sampler(_far3(n), [0.0, 0.5, 1.5], D4), 3),
    # far field, moving materials, a different u2 pair, 20 lags (two lag chunks)
    "far3_moving_pair": (lambda n: G.corr_sampler(_far3(n), np.linspace(0, 3, 20), D4, U4,
                                                  G.perturbed(_far3(n)["illumination"], 1, 0.03),
                                                  G.perturbed(_far3(n)["view"], 2, 0.03)), 3),
    # point sources, anisotropic scattering: the generic single-scattering kernel
    "point3_pair": (lambda n: G.corr_sampler(_pt3(n), [0.0, 0.7, 2.0], D4, U4,
                                             G.perturbed(_pt3(n)["illumination"], 3, 0.3),
                                             G.perturbed(_pt3(n)["view"], 4, 0.3)), 3),
    # point sources, isotropic scattering: scalar Fs (no single-scattering decorrelation)
    "point3_iso_scalar": (lambda n: G.corr_sampler(_pt3_iso(n), [0.0, 1.0], D4, U4), 3),
    # 2D point sources and a moving pair
    "point2_pair": (lambda n: G.corr_sampler(G.hetero2_points(n), [0.0, 0.4, 1.2], D4, U4,
                                             G.perturbed(G.hetero2_points(n)["illumination"], 5, 0.2),
                                             G.perturbed(G.hetero2_points(n)["view"], 6, 0.2)), 2),
    # tabular phase functions (complex amplitudes) in the correlation terms
    "far3_tabular_pair": (lambda n: G.corr_sampler(G.with_tabular(_far3(n), 3), [0.0, 1.0], D4, U4,
                                                   G.perturbed(_far3(n)["illumination"], 7, 0.02),
                                                   G.perturbed(_far3(n)["view"], 8, 0.02)), 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_terms_match_oracle(case):
    make, dims = CASES[case]
    cfg = make(128)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    print("case", case, "vertices", int(d["off"][-1]), flush=True)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    assert np.count_nonzero(d["k"] >= 2) > 10
    assert _rel(d["a"], A) < 1e-10
    assert _rel(d["b"], B) < 1e-10
    assert _rel(d["u"], U) < 1e-10


@pytest.mark.parametrize("case", ["far3_moving_pair", "point3_pair", "point2_pair"])
def test_fp32_terms(case):
    """FP32 terms: batch correlation within 2e-4 of the FP64 restatement on the same vertices."""
    make, dims = CASES[case]
    cfg = make(128)
    cfg["precision"] = {"terms": "fp32", "contraction": "fp32"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    assert _rel(d["a"], A) < 1e-4
    assert _rel(d["b"], B) < 1e-4
    assert _rel(d["u"], U) < 2e-4


@pytest.mark.parametrize("case,dims", [("far3_moving_pair", 3), ("point2_pair", 2)])
def test_transport_draws_from_both_source_sets(case, dims):
    cfg = CASES[case][0](512)
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    o, same_k, match = _transport_match(cfg, dims, d, n, tol=2e-3)
    assert np.mean(match) >= 0.95, (np.mean(same_k), np.mean(match))


def test_mcx_known_answer_on_gpu():
    """examples/mcxValidation3D.m with the TemporalCorrelationSampler: C = u / iter against mcx_res
    over 8 renderings of 200000 paths (level within 4 standard errors, shape C(t) / C(0) within
    3 %)."""
    r = json.load(open(os.path.join(HERE, "golden", "mcx_validation.json")))
    cfg = W.mcx_validation(r["t"], r["D"], r["MFP"], r["R"], r["h"], r["dims"], r["viSep"], batch=20000,
                           full_iteration=10, seed=17)
    cfg = G.corr_sampler(cfg, cfg["sampler"]["t"], cfg["sampler"]["D"])
    cfg["simulation"]["renderingsNum"] = 8
    u, it = sstmc3D(cfg)
    res = u[0, 0, :, :].real / it[None, :]          # [t, rendering]
    mcx = np.array(r["mcx_res"])
    m = res.mean(1)
    se = res.std(1, ddof=1) / math.sqrt(res.shape[1])
    print("level", m[0] / mcx[0], "shape end", (m[-1] / m[0]) / (mcx[-1] / mcx[0]))
    assert np.all(np.abs(m - mcx) < 4 * se)
    shape = (m / m[0]) / (mcx / mcx[0])
    assert np.abs(shape - 1).max() < 0.03, shape


@pytest.mark.parametrize("motion", ["Linear", "Brownian-Linear"])
def test_spatio_temporal_example_matches_oracle(motion):
    """examples/spatioTemporalCorrelation2D.m (11 pixels, its tabular amplitude from
    inPhasefunction.mat): FP64 terms and batch correlation against the restatement."""
    amp = np.load(os.path.join(HERE, "golden", "in_phase_function.npz"))["amp"]
    cfg = W.spatio_temporal_2d(amp, pixels=11, motion=motion, batch=512, full_iteration=1, seed=3)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(cfg, 2) as tr:
        d = tr.debug_batch(0, 512)
    A, B, S, U = _oracle_u(cfg, 2, d, d["a"], d["b"])
    assert _rel(d["a"], A) < 1e-10 and _rel(d["b"], B) < 1e-10
    assert _rel(d["u"], U) < 1e-10
