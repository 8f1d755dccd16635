"""GPU parity: the HIP path against the FP64 CPU restatement (oracle/) on identical inputs.

Three layers, so that a mismatch names the stage that caused it:
  1. transport: vertex counts and positions from the FP32 GPU transport vs the FP64 oracle
     transport on the same counter-based random numbers (paths whose discrete decisions flip
     near a threshold are counted, not hidden);
  2. terms: NEE terms computed by the GPU from its own vertices vs the oracle's FP64 terms from
     the same vertices (amplitude and phase tolerances stated per test);
  3. contraction: the batch field u from the GPU vs the sum of the oracle's per-path
     contributions on the same vertices.
"""
import numpy as np
import pytest

import oracle_lib as O
import gpu_cases as G
from temporalspecklescode_amd import Tracer

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _oracle_u(cfg, dims, d, a, b, s_unused=None):
    """Batch field from oracle terms on the GPU's vertices: sum_p S_p + A_p (x) B_p. For the
    correlation output A and B already hold the per-path lag products; the single-scattering
    field S_p[t] is turned into S_p[0] conj(S_p[t]) here."""
    A, B, S = O.terms_from_vertices(cfg, dims, d["k"], d["off"], d["pos"].astype(np.float64), d["mat"],
                                    d["rho"].astype(np.float64), d["xt"], d["len"])
    if cfg.get("tracer", {}).get("output") == "correlation":
        S = S[:, :1] * np.conj(S)
    U = S.sum(axis=0).transpose(2, 1, 0) + np.einsum("ptl,ptv->lvt", A, B)
    return A, B, S, U


def _transport_match(cfg, dims, d, n, tol, terms=True):
    o = O.paths(cfg, dims, 0, n, terms=terms)
    same_k = d["k"] == o["k"]
    match = np.zeros(n, bool)
    for i in np.nonzero(same_k)[0]:
        g0, g1 = int(d["off"][i]), int(d["off"][i + 1])
        o0, o1 = int(o["off"][i]), int(o["off"][i + 1])
        if np.max(np.abs(d["pos"][g0:g1].astype(np.float64) - o["pos"][o0:o1]), initial=0.0) < tol:
            match[i] = True
    return o, same_k, match


@pytest.mark.parametrize("case,dims,min_match", [
    ("c1", 2, 0.99), ("hetero_ff", 3, 0.95), ("hetero_ps", 3, 0.95), ("frames", 3, 0.99)])
def test_transport_matches_oracle(case, dims, min_match):
    cfg = {"c1": G.c1_small(1024), "hetero_ff": G.hetero3_small(512), "hetero_ps": G.hetero3_small(512, point=True),
           "frames": G.slab_frames(512)}[case]
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    print("case", case, "paths", n, "vertices", int(d["off"][-1]), flush=True)
    o, same_k, match = _transport_match(cfg, dims, d, n, tol=2e-3)
    assert np.mean(d["misses"] == o["misses"]) > 0.99
    assert np.mean(match) >= min_match, (np.mean(same_k), np.mean(match))
    # rho and frame displacements come from their own streams: identical where the path matches
    i = int(np.nonzero(match)[0][0])
    g0, o0 = int(d["off"][i]), int(o["off"][i])
    kk = int(d["k"][i])
    assert np.allclose(d["rho"][g0:g0 + kk], o["rho"][o0:o0 + kk], atol=1e-7)
    assert np.allclose(d["xt"][g0:g0 + kk], o["xt"][o0:o0 + kk], atol=2e-3)


FP64_CASES = {
    "c1": (lambda: G.c1_small(256), 2),
    "hetero_ff": (lambda: G.hetero3_small(128), 3),
    "hetero_ps": (lambda: G.hetero3_small(128, point=True), 3),
    "frames": (lambda: G.slab_frames(128), 3),
    "frames_corr": (lambda: G.slab_frames(128, output="correlation"), 3),
    "frames_linear_motion": (lambda: G.slab_frames(128, U=0.02), 3),
    "points_2d_nonuniform": (lambda: G.hetero2_points(128), 2),
    "points_2d_area": (lambda: G.hetero2_points(128, area=True), 2),
    "moving_ff_points": (lambda: G.moving_ff_points(128), 3),
    "moving_no_temporal_amplitude": (lambda: G.moving_ff_points(128, temporal_amp=False), 3),
}


@pytest.mark.parametrize("case", sorted(FP64_CASES))
def test_fp64_terms_match_oracle_on_gpu_vertices(case):
    make, dims = FP64_CASES[case]
    cfg = make()
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    print("case", case, "vertices", int(d["off"][-1]), flush=True)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    assert _rel(d["a"], A) < 1e-10
    assert _rel(d["b"], B) < 1e-10
    assert _rel(d["u"], U) < 1e-10


@pytest.mark.parametrize("case,dims", [("c1", 2), ("hetero_ff", 3), ("hetero_ps", 3), ("frames", 3)])
def test_fp32_terms_amplitude_and_phase(case, dims):
    """FP32 NEE (walks, exponentials, phase functions) with FP64 phase. Single terms (A, a, b):
    per-path amplitude relative error <= 1e-5 (Frobenius over the path's L or V terms) and phase
    error <= 1e-4 rad on every term within 1e-3 of the path's largest. B_v = sum_k G_k(v) adds
    K-1 such terms with random phases; their errors add while the sum partly cancels
    (E|sum| ~ sqrt(K-1) |G|), so B is held to a complex relative error <= 5e-5 per path."""
    cfg = {"c1": G.c1_small(256), "hetero_ff": G.hetero3_small(128), "hetero_ps": G.hetero3_small(128, point=True),
           "frames": G.slab_frames(128)}[case]
    cfg["precision"] = {"terms": "fp32", "contraction": "fp32"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    for name, g, r in (("A", d["a"], A), ("sa", d["sa"], None), ("B", d["b"], B)):
        if r is None:
            continue
        for p in range(n):
            ref = r[p].ravel()
            got = g[p].ravel()
            nr = np.linalg.norm(ref)
            if nr == 0:
                assert np.linalg.norm(got) == 0
                continue
            if name == "B":
                err = np.linalg.norm(got - ref) / nr
                assert err < 5e-5, (name, p, err)
                continue
            amp = np.linalg.norm(np.abs(got) - np.abs(ref)) / nr
            big = np.abs(ref) > 1e-3 * np.abs(ref).max()
            dph = np.max(np.abs(np.angle(got[big] / ref[big])))
            assert amp < 1e-5 and dph < 1e-4, (name, p, amp, dph)
    assert _rel(d["u"], U) < 1e-4


@pytest.mark.parametrize("kind,tol", [("bf16x3", 1e-4), ("fp32", 1e-5), ("fp64", 1e-12)])
def test_contraction_against_fp64_gemm(kind, tol):
    cfg = G.hetero3_small(64)
    L, V = 4, 4
    cfg["precision"] = {"terms": "fp64" if kind == "fp64" else "fp32", "contraction": kind}
    rng = np.random.default_rng(1)
    for (P, L, V, T) in ((1000, 37, 45, 1), (4096, 130, 129, 2)):
        c = G.slab_frames(64, nl=L, nv=V, nt=T)
        c["precision"] = cfg["precision"]
        a = rng.normal(size=(P, T, L)) + 1j * rng.normal(size=(P, T, L))
        b = (rng.normal(size=(P, T, V)) + 1j * rng.normal(size=(P, T, V))) * np.exp(rng.normal(size=(P, 1, 1)))
        with Tracer(c, 3) as tr:
            u = tr.debug_contract(a, b)
        ref = np.einsum("ptl,ptv->lvt", a, b)
        print(kind, P, L, V, T, "rel err", _rel(u, ref), flush=True)
        assert _rel(u, ref) < tol


def test_render_deterministic_and_iter():
    cfg = G.c1_small(4096)
    cfg["simulation"]["batchSize"] = 1024
    cfg["simulation"]["fullIteration"] = 4
    with Tracer(cfg, 2) as tr:
        u1, it1 = tr.trace()
        u2, it2 = tr.trace()
        miss = tr.zero_contribution_paths
    assert np.array_equal(u1, u2) and it1 == it2
    assert it1 == 4096 + miss
    # the oracle on the same counter-based paths: same miss count
    uo, ito, _ = O.render(cfg, 2)
    assert abs(ito - it1) <= 0.01 * 4096


@pytest.mark.parametrize("case", ["moving_points_amp", "moving_points", "tabular_points_2d"])
def test_correlation_output_with_per_path_single_scattering(case):
    """Correlation output where the single-scattering amplitude depends on the first vertex (point
    sources with HG or tabular scattering): sum_p a b f(0) conj f(t), with f(t) at the displaced
    first vertex when the amplitude follows the scatterers."""
    if case == "tabular_points_2d":
        cfg, dims = G.with_tabular(G.hetero2_points(128), 2), 2
        cfg["sampler"] = {"type": "TemporalSampler", "t": np.array([0.0, 1.0, 2.0]), "D": np.full(4, 2e-3),
                          "U": {"x": np.zeros(4), "y": np.zeros(4)}, "isTemporalAmplitude": True}
    else:
        cfg, dims = G.moving_ff_points(128, temporal_amp=case == "moving_points_amp"), 3
    cfg["tracer"] = {"type": "nee", "output": "correlation"}
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    print("case", case, "vertices", int(d["off"][-1]), flush=True)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    assert np.linalg.norm(S) > 0
    assert _rel(d["u"], U) < 1e-10
