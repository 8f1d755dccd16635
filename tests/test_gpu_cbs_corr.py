"""Coherent backscattering under the TemporalCorrelationSampler on the GPU against the FP64 CPU
restatement on the GPU's own vertices (NextEventEstimation.cuh:1138-1391): CBS04's halves Ar / Br,
and the batch correlation S + (A^T B + Br^T Ar + CBS02 + CBS03) / 2 with the mixed pairings formed
per (l, v) on the device; far-field and point sources, moving materials, a different u2 pair, lag
chunks; and the reference's known answer, examples/mudiffCBSValidation2D.m:66-90, on the GPU."""
import math
import os

import numpy as np
import pytest

import gpu_cases as G
import oracle_lib as O
from temporalspecklescode_amd import Tracer, sstmc2D
from temporalspecklescode_amd import workloads as W
from test_gpu_parity import _rel

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

D4 = [0.0, 0.01, 0.02, 0.005]
U4 = [[0, 0, 0], [0.05, 0.0, 0.02], [-0.03, 0.04, 0.0], [0.0, 0.0, 0.0]]


def _cbs(cfg):
    cfg["tracer"] = dict(cfg["tracer"], isCBS=True)
    return cfg


CASES = {
    # far field, Brownian only, u2 = u1
    "far3_static_same": (lambda n: _cbs(G.corr_sampler(G.hetero3_small(n), [0.0, 0.5, 1.5], D4)), 3),
    # far field, moving materials, a different u2 pair, 20 lags (two lag chunks)
    "far3_moving_pair": (lambda n: _cbs(G.corr_sampler(G.hetero3_small(n), np.linspace(0, 3, 20), D4, U4,
                                                       G.perturbed(G.hetero3_small(n)["illumination"], 1, 0.03),
                                                       G.perturbed(G.hetero3_small(n)["view"], 2, 0.03))), 3),
    # point sources, a moving pair
    "point3_pair": (lambda n: _cbs(G.corr_sampler(G.hetero3_small(n, point=True), [0.0, 0.7, 2.0], D4, U4,
                                                  G.perturbed(G.hetero3_small(n, point=True)["illumination"], 3, 0.3),
                                                  G.perturbed(G.hetero3_small(n, point=True)["view"], 4, 0.3))), 3),
    # reflection geometry (views = negated illuminations): the mixed pairings carry the cone
    "backscatter_far": (lambda n: _cbs(G.corr_sampler(G.cbs_backscatter(n), [0.0, 0.5, 1.5], D4)), 3),
    "backscatter_point_moving": (lambda n: _cbs(G.corr_sampler(G.cbs_backscatter(n, point=True), [0.0, 1.0], D4,
                                                               U4)), 3),
    # 2D point sources
    "point2_pair": (lambda n: _cbs(G.corr_sampler(G.hetero2_points(n), [0.0, 0.4, 1.2], D4, U4,
                                                  G.perturbed(G.hetero2_points(n)["illumination"], 5, 0.2),
                                                  G.perturbed(G.hetero2_points(n)["view"], 6, 0.2))), 2),
}


def _oracle(cfg, dims, d):
    args = (d["k"], d["off"], d["pos"].astype(np.float64), d["mat"], d["rho"].astype(np.float64), d["xt"], d["len"])
    A, B, S = O.terms_from_vertices(cfg, dims, *args)
    Ar, Br = O.cbs_terms_from_vertices(cfg, dims, *args)
    X = O.cbs_cross_from_vertices(cfg, dims, *args)
    U = S.sum(axis=0).transpose(2, 1, 0) + 0.5 * (np.einsum("ptl,ptv->lvt", A, B) +
                                                  np.einsum("ptl,ptv->lvt", Br, Ar) + X)
    return A, B, Ar, Br, X, U


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_terms_and_correlation_match_oracle(case):
    make, dims = CASES[case]
    cfg = make(128)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
        ar, br = tr.debug_cbs(n)
    print("case", case, "paths", n, "vertices", int(d["off"][-1]), flush=True)
    A, B, Ar, Br, X, U = _oracle(cfg, dims, d)
    assert np.count_nonzero(d["k"] >= 2) > 10
    ea, eb, ear, ebr, eu = (_rel(d["a"], A), _rel(d["b"], B), _rel(ar, Ar), _rel(br, Br), _rel(d["u"], U))
    # the device's mixed pairings, isolated from its batch correlation
    xg = 2 * (d["u"] - (U - 0.5 * X)) - (np.einsum("ptl,ptv->lvt", d["a"], d["b"]) +
                                         np.einsum("ptl,ptv->lvt", br, ar) - np.einsum("ptl,ptv->lvt", A, B) -
                                         np.einsum("ptl,ptv->lvt", Br, Ar))
    ex = _rel(xg, X)
    print("rel err A %.2e B %.2e Ar %.2e Br %.2e u %.2e mixed %.2e; mixed pairings %.2e of u" %
          (ea, eb, ear, ebr, eu, ex, np.linalg.norm(X) / np.linalg.norm(U)), flush=True)
    assert max(ea, eb, ear, ebr, eu) < 1e-10
    # (isolated through the batch sum: also allow that sum's rounding when the pairings are small)
    assert np.linalg.norm(xg - X) < 1e-8 * np.linalg.norm(X) + 1e-12 * np.linalg.norm(U)
    if case.startswith("backscatter"):
        assert np.linalg.norm(X) > 0.05 * np.linalg.norm(U)


@pytest.mark.parametrize("case", ["far3_moving_pair", "point2_pair"])
def test_fp32_terms(case):
    """FP32 terms and contraction: the batch correlation within 2e-4 of the FP64 restatement."""
    make, dims = CASES[case]
    cfg = make(128)
    cfg["precision"] = {"terms": "fp32", "contraction": "fp32"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    _, _, _, _, _, U = _oracle(cfg, dims, d)
    err = _rel(d["u"], U)
    print("case", case, "fp32 rel err %.2e" % err, flush=True)
    assert err < 2e-4


def test_render_deterministic_with_pipeline():
    """A whole rendering (several batches through the three-stage pipeline): run twice, the same bits."""
    cfg = CASES["far3_moving_pair"][0](1024)
    cfg["simulation"]["batchSize"] = 256
    from temporalspecklescode_amd import render
    u1, it1 = render(cfg, 3)
    u2, it2 = render(cfg, 3)
    assert np.array_equal(u1, u2) and np.array_equal(it1, it2)
    assert np.all(np.isfinite(u1)) and np.linalg.norm(u1) > 0


def test_mudiff_correlation_cbs_known_answer_on_gpu():
    """examples/mudiffCBSValidation2D.m:66-90: the TemporalCorrelationSampler with isCBS over 10
    renderings of 10 x 1000 paths against mudiffRes at 721 view angles, cone included, each
    comparison allowing the estimator's standard error and mudiffRes' own scatter (its rms
    neighbour difference over sqrt 2)."""
    fx = dict(np.load(os.path.join(HERE, "golden", "mudiff_cbs.npz"), allow_pickle=False))
    ref = fx["mudiff"]
    back = int(np.argmin(np.abs(fx["deg"] - math.pi)))
    ref_sd = np.sqrt(np.mean((np.diff(ref) / ref[1:]) ** 2) / 2) * ref
    cfg = W.mudiff_cbs(fx["amp"], fx["deg"], fx["x"], fx["y"], float(fx["mfp"]), batch=1000, renderings=10,
                       cbs=True, seed=3)
    cfg["simulation"]["fullIteration"] = 10
    cfg = G.corr_sampler(cfg, [0.0], [0.0, 0.0])
    cfg["tracer"] = {"type": "nee", "isCBS": True}
    u, it = sstmc2D(cfg)                        # u: (1, 721, 1, 10)
    x = u[0, :, 0, :].real / it[None, :]
    mean, se = x.mean(axis=1), x.std(axis=1, ddof=1) / math.sqrt(x.shape[1])
    z = (mean - ref) / np.sqrt(se ** 2 + ref_sd ** 2)
    print("mean ratio", mean.mean() / ref.mean(), "max |z|", np.abs(z).max(), "cone", mean[back], ref[back],
          flush=True)
    assert abs(mean.mean() / ref.mean() - 1) < 0.02
    assert np.mean(np.abs(z) > 4) <= 0.01
    assert abs(z[back]) < 4 and mean[back] > 1.4 * np.median(mean)
