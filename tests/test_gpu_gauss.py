"""GaussianBeam sources (GaussianBeam.cuh, 3D) on the GPU against the FP64 CPU restatement, on the
confocal example's geometry (four vMF-mixture cells, beams focused along +z): the beam-profile
first point and the lobe-pair second direction in transport, the first- and second-point
probabilities, the beams' complex f at the first and last vertices, the fs single-scattering
kernel, and moving frames with the amplitude following the scatterers. Terms are compared on the
GPU's own vertices; the rendered field through its confocal gating property."""
import numpy as np
import pytest

import oracle_lib as O
from temporalspecklescode_amd import Tracer
from temporalspecklescode_amd import workloads as W
from test_gpu_parity import _oracle_u, _rel, _transport_match

pytestmark = pytest.mark.gpu


def _small(batch=128, frames=False, seed=1, shift=1.0):
    cfg = W.confocal(focal_x=np.arange(-4.0, 5.0, 2.0), focal_z=np.array([4.0, 12.0]), tau_x=(0.0, shift),
                     batch=batch, full_iteration=1, seed=seed)
    if frames:
        cfg["sampler"] = {"type": "TemporalSampler", "t": np.array([0.0, 1.0, 2.5]),
                          "D": np.array([0.0, 1e-3, 0.0, 4e-3, 2e-3]),
                          "U": {"x": np.array([0.0, 0.02, -0.01, 0.0, 0.01]), "y": np.zeros(5),
                                "z": np.array([0.0, 0.0, 0.03, 0.01, 0.0])},
                          "isTemporalAmplitude": True}
    return cfg


def _tilted(cfg, seed=7):
    """View beams along individually tilted unit directions (the general fs kernel; with one shared
    view direction the per-illumination part of fs is formed once per path and shared)."""
    n = len(cfg["view"]["focalPoint"]["x"])
    d = np.random.default_rng(seed).normal(size=(n, 3)) * [0.15, 0.15, 0.0] + [0.0, 0.0, 1.0]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cfg["view"]["focalDirection"] = {"x": d[:, 0], "y": d[:, 1], "z": d[:, 2]}
    return cfg


@pytest.mark.parametrize("frames,tilted", [(False, False), (True, False), (False, True)])
def test_fp64_terms_match_oracle(frames, tilted):
    cfg = _small(128, frames)
    if tilted:
        cfg = _tilted(cfg)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, n)
    print("frames", frames, "tilted", tilted, "vertices", int(d["off"][-1]), flush=True)
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    assert np.linalg.norm(A) > 0 and np.linalg.norm(B) > 0
    assert _rel(d["a"], A) < 1e-9
    assert _rel(d["b"], B) < 1e-9
    assert _rel(d["u"], U) < 1e-9


@pytest.mark.parametrize("tilted", [False, True])
def test_fp32_terms(tilted):
    """FP32 walks and products over FP64 beam functions: batch field within 1e-4."""
    cfg = _small(128)
    if tilted:
        cfg = _tilted(cfg)
    cfg["precision"] = {"terms": "fp32", "contraction": "fp32"}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, n)
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    assert _rel(d["a"], A) < 1e-4
    assert _rel(d["b"], B) < 1e-4
    assert _rel(d["u"], U) < 1e-4


def test_transport_samples_the_beam():
    cfg = _small(512, seed=3)
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, n)
    o, same_k, match = _transport_match(cfg, 3, d, n, tol=2e-3)
    print("same k", np.mean(same_k), "match", np.mean(match), flush=True)
    assert np.mean(d["misses"] == o["misses"]) > 0.99
    assert np.mean(match) >= 0.95, (np.mean(same_k), np.mean(match))


def test_render_gating_and_iter():
    """The rendered field peaks where the view focus coincides with the illumination's (the
    confocal example's gating); two renders are identical."""
    cfg = _small(1000, shift=2.0)
    cfg["simulation"]["fullIteration"] = 2
    with Tracer(cfg, 3) as tr:
        u1, it = tr.trace(seed=5)
        u2, _ = tr.trace(seed=5)
    assert it == 2000 and np.array_equal(u1, u2) and np.isfinite(u1).all()
    u = u1[:, :, 0]
    L = u.shape[0]
    same = np.abs([u[i, i] for i in range(L)])
    shifted = np.abs([u[i, L + i] for i in range(L)])
    assert same.mean() > 3 * shifted.mean()


def test_wide_grid_fp64_matches_oracle():
    """Beams up to 100 wavelengths from the first vertex: large imaginary throughputs, whose shared
    fs factors must keep each exponent whole (finite field, FP64 parity)."""
    cfg = W.confocal(focal_x=np.array([-50.0, -25.0, 0.0, 25.0, 50.0]), focal_z=np.array([2.0, 10.0, 18.0]),
                     tau_x=(-1.0, 0.0, 1.0), batch=512, full_iteration=1, seed=11)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, 512)
    assert np.isfinite(d["u"]).all()
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    assert _rel(d["u"], U) < 1e-9


@pytest.mark.parametrize("amp", [False, True])
def test_correlation_output(amp):
    """Correlation output over moving frames: view sums G(0) conj G(t) with the beams' complex f,
    single scattering fs(0) conj fs(t) (fs at the displaced first vertex when the amplitude follows
    the scatterers)."""
    cfg = _small(128, frames=True)
    cfg["sampler"]["isTemporalAmplitude"] = amp
    cfg["tracer"] = {"type": "nee", "output": "correlation"}
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, 128)
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    assert np.linalg.norm(S) > 0 and np.linalg.norm(B) > 0
    assert _rel(d["u"], U) < 1e-9


@pytest.mark.parametrize("backscattering", [True, False])
def test_transmission_matrix_example(backscattering):
    """The reference's transmissionMatrixRendering.m geometry (5 x 5 pixels here): isotropic vMF
    lobes everywhere, so the second direction is uniform and its density 1 / (4 pi); focal points on
    the z = 0 face; views looking back along -z or through the slab; five Brownian frames."""
    cfg = W.gauss_transmission_matrix(pixels=5, dp=2.0, backscattering=backscattering, batch=256,
                                      full_iteration=1, seed=4)
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, 256)
    o, same_k, match = _transport_match(cfg, 3, d, 256, tol=2e-3)
    print("backscattering", backscattering, "vertices", int(d["off"][-1]), "match", np.mean(match), flush=True)
    assert np.mean(match) >= 0.95
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    assert np.linalg.norm(S) > 0 and np.linalg.norm(B) > 0
    assert _rel(d["u"], U) < 1e-9


@pytest.mark.parametrize("case", ["moving_amplitude_fixed", "static_ten_frames"])
def test_shared_direction_fs_frames(case):
    """One view direction for every beam and field output: the shared-direction fs kernel with
    frames sharing one fs evaluation (the first vertex does not move with the frame), over three
    moving frames with isTemporalAmplitude false, and over ten static frames (an 8-frame chunk and a
    partial one of 2)."""
    cfg = _small(128)
    if case == "moving_amplitude_fixed":
        cfg["sampler"] = {"type": "TemporalSampler", "t": np.array([0.0, 1.0, 2.5]),
                          "D": np.array([0.0, 1e-3, 0.0, 4e-3, 2e-3]),
                          "U": {"x": np.array([0.0, 0.02, -0.01, 0.0, 0.01]), "y": np.zeros(5),
                                "z": np.array([0.0, 0.0, 0.03, 0.01, 0.0])},
                          "isTemporalAmplitude": False}
    else:
        cfg["sampler"] = {"type": "TemporalSampler", "t": np.arange(10.0), "D": np.zeros(5),
                          "U": {"x": np.zeros(5), "y": np.zeros(5), "z": np.zeros(5)}}
    cfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(cfg, 3) as tr:
        d = tr.debug_batch(0, 128)
    A, B, S, U = _oracle_u(cfg, 3, d, d["a"], d["b"])
    nt = U.shape[2]
    print("case", case, "frames", nt, "single-scattering share %.3f" % (np.linalg.norm(S.sum(0)) / np.linalg.norm(U)),
          flush=True)
    assert np.linalg.norm(S) > 0
    assert _rel(d["u"], U) < 1e-9
    for t in range(nt):   # every frame, the partial chunk's included
        assert _rel(d["u"][:, :, t], U[:, :, t]) < 1e-9, t
