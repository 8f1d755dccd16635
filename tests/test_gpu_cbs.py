"""Coherent backscattering on the GPU against the FP64 CPU restatement, on the GPU's own vertices:
the reversed-path terms Ar (view connected at x1) and Br (illumination connected at x_k >= 2) and
the batch field S + (A^T B + Br^T Ar) / sqrt(2) (NextEventEstimation.cuh:1212-1235,1385-1391,
1432-1454)."""
import numpy as np
import pytest

import oracle_lib as O
import gpu_cases as G
from temporalspecklescode_amd import Tracer

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _oracle(cfg, d):
    args = (d["k"], d["off"], d["pos"].astype(np.float64), d["mat"], d["rho"].astype(np.float64), d["xt"], d["len"])
    A, B, S = O.terms_from_vertices(cfg, 3, *args)
    Ar, Br = O.cbs_terms_from_vertices(cfg, 3, *args)
    U = S.sum(axis=0).transpose(2, 1, 0) + (np.einsum("ptl,ptv->lvt", A, B) +
                                            np.einsum("ptl,ptv->lvt", Br, Ar)) / np.sqrt(2.0)
    return A, B, Ar, Br, U


@pytest.mark.parametrize("point", [False, True])
def test_cbs_fp64_terms_and_field_match_oracle(point):
    cfg = G.cbs_backscatter(paths=192, point=point)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, n)
        ar, br = tr.debug_cbs(n)
    A, B, Ar, Br, U = _oracle(cfg, d)
    print("point", point, "multi paths", int((d["k"] >= 2).sum()), flush=True)
    assert _rel(d["a"], A) < 1e-10 and _rel(d["b"], B) < 1e-10
    assert _rel(ar, Ar) < 1e-10, _rel(ar, Ar)
    assert _rel(br, Br) < 1e-10, _rel(br, Br)
    assert _rel(d["u"], U) < 1e-10, _rel(d["u"], U)


def test_cbs_reversed_terms_equal_direct_terms_under_negated_views():
    # reflection geometry with view v = -illumination v: Ar = A and Br = B index for index, so
    # the multiple-scattering field is (M + M^T) / sqrt(2) and the exact-backscatter diagonal is
    # enhanced by sqrt(2) in amplitude (2 in intensity)
    cfg = G.cbs_backscatter(paths=256)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, n)
        ar, br = tr.debug_cbs(n)
    assert _rel(ar, d["a"]) < 1e-13 and _rel(br, d["b"]) < 1e-13
    M = np.einsum("ptl,ptv->lvt", d["a"], d["b"])
    S = d["u"] - (M + np.transpose(M, (1, 0, 2))) / np.sqrt(2.0)
    plain = dict(cfg, tracer={"type": "nee"})
    with Tracer(plain, 3) as tp:
        d0 = tp.debug_batch(0, n)
    assert _rel(d0["u"], S + M) < 1e-10


def test_cbs_fp32_terms_and_mixed_field():
    cfg = G.cbs_backscatter(paths=256)
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:                      # FP32 terms, bf16x3 contraction
        d = tr.debug_batch(0, n)
        ar, br = tr.debug_cbs(n)
    A, B, Ar, Br, U = _oracle(cfg, d)
    multi = d["k"] >= 2
    for p in np.nonzero(multi)[0]:
        nr = np.linalg.norm(Ar[p])
        amp = np.linalg.norm(np.abs(ar[p]) - np.abs(Ar[p])) / nr
        big = np.abs(Ar[p]) > 1e-3 * np.abs(Ar[p]).max()
        dph = np.max(np.abs(np.angle(ar[p][big] / Ar[p][big])))
        assert amp < 1e-5 and dph < 1e-4, (p, amp, dph)
        assert np.linalg.norm(br[p] - Br[p]) / np.linalg.norm(Br[p]) < 5e-5, p
    assert _rel(d["u"], U) < 1e-4


def test_cbs_render_runs_and_scales():
    # whole render through the C ABI with CBS: deterministic, and equal to the debug path's field
    cfg = G.cbs_backscatter(paths=512)
    cfg["simulation"]["batchSize"] = 256
    from temporalspecklescode_amd import render
    u1, it1 = render(cfg, 3)
    u2, it2 = render(cfg, 3)
    assert np.array_equal(u1, u2) and np.array_equal(it1, it2)
    assert np.all(np.isfinite(u1)) and np.linalg.norm(u1) > 0
