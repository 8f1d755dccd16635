"""2D GaussianBeam sources (GaussianBeam.cuh, DIMS == 2) in the CPU restatement.

Pinned independently of the restatement's own formulas:
  * log I_0(z) for complex z against scipy.special.iv (series and large-argument branches, near
    the imaginary axis too);
  * f (fGaussianBeam, :876-878) is the circle integral sum_j exp(apNorm + C_j)
    int_0^2pi exp((gamma_j w0 + T) . w(theta)) dtheta of the vMF lobes against the beam's complex
    throughput T = gamma_a d + i k (x0 - focal), apNorm = -log 2 pi - log I_0(gamma_a): trapezoid
    quadrature (spectrally accurate for the periodic integrand), illumination and view;
  * fs (fsGaussianBeam, :989-991) of a lobe with gamma = 0 factorises into the two beams' circle
    integrals, exactly;
  * the first point is drawn as from a point source at the focal point (sampleFirstBeam, :754-773,
    then the medium's sample): the same vertices as a PointSource configuration with the focal
    points as locations.
Parity unpinned beyond the reference's text: fs for lobes with gamma != 0, the reference's closed
form of the double integral, compared here with the integral itself to within the approximation's
size (reported).
"""
import math

import numpy as np
import pytest
from scipy import special

import oracle_lib as O

LOBES = {"mixtureMu": [0.0, 3.0, 12.0], "mixtureC": [-1.0, -2.0, -6.0], "mixtureAlpha": [0.5, 0.3, 0.2]}


def _cfg(lobes=LOBES, aperture=0.4, lam=1.0, focal=((0.0, 8.0),), view=((0.5, 8.0),), batch=64):
    def beams(pts, d=(0.0, 1.0)):
        p = np.asarray(pts, np.float64)
        return {"type": "GaussianBeam", "focalPoint": {"x": p[:, 0], "y": p[:, 1]},
                "focalDirection": {"x": np.full(len(p), d[0]), "y": np.full(len(p), d[1])}, "aperture": aperture}

    return {
        "simulation": {"batchSize": batch, "seed": 5, "lambda": lam},
        "medium": {"type": "Heterogeneous", "xAxis": np.array([-10.0, 10.0]), "yAxis": np.array([0.0, 16.0]),
                   "materialGrid": np.ones((1, 1), np.uint32), "sigs": [0.0, 0.25], "siga": [0.0, 0.01]},
        "scattering": {"type": ["vMF"] * 2, "amplitudeFunction": [dict(lobes)] * 2},
        "illumination": beams(focal),
        "view": beams(view),
        "sampler": {"type": "TemporalSampler", "t": np.zeros(1), "D": np.zeros(2),
                    "U": {"x": np.zeros(2), "y": np.zeros(2)}},
        "tracer": {"type": "nee"},
    }


def _circle_integral(v, n=4096):
    """int_0^2pi exp(v . w(theta)) dtheta for a complex 2-vector v (trapezoid, spectral)."""
    th = np.arange(n) * 2 * np.pi / n
    w = np.stack([np.cos(th), np.sin(th)], axis=1)
    return np.sum(np.exp(w @ v)) * 2 * np.pi / n


def _lognorm2(x):
    return -math.log(2 * math.pi) - (math.log(special.ive(0, x)) + x)


def test_log_i0_matches_library():
    zs = []
    for r in (0.0, 0.3, 2.0, 7.5, 11.9, 12.1, 20.0, 45.0, 150.0):
        for a in np.linspace(-math.pi / 2, math.pi / 2, 9):
            zs.append(r * complex(math.cos(a), math.sin(a)))
    for z in zs:
        got = O.log_i0(z)
        ref = special.iv(0, z)
        if abs(z) > 600:
            continue
        # compare I_0 itself (the log's imaginary part is defined modulo 2 pi); the series branch
        # loses about e^|z| 2^-53 to cancellation near the imaginary axis, where |I_0| ~ |J_0| is
        # small (bounded against the largest term, e^|z|), and the expansion's least term at the
        # switch, |z| = 12, is about 1e-11 relative
        err = abs(np.exp(got) - ref)
        assert err <= 5e-11 * max(abs(ref), 1.0) + 1e-15 * math.exp(abs(z)), (z, np.exp(got), ref)
    # large real arguments stay finite in log form where I_0 overflows
    assert abs(O.log_i0(1000.0) - (1000.0 - 0.5 * math.log(2 * math.pi * 1000.0) + math.log1p(1 / 8000.0 + 9 / 128e6))) < 1e-9


@pytest.mark.parametrize("view", [0, 1])
def test_f_is_the_circle_integral(view):
    cfg = _cfg()
    lam, ap = 1.0, 0.4
    k = 2 * math.pi / lam * (-1 if view else 1)
    ga = 1 / ap ** 2
    focal = np.array([0.0, 8.0]) if not view else np.array([0.5, 8.0])
    rng = np.random.default_rng(3)
    for _ in range(5):
        x0 = focal + rng.uniform(-0.4, 0.4, 2)
        p2 = x0 + rng.normal(size=2) * 3
        w0 = (p2 - x0) if not view else (x0 - p2)
        w0 /= np.linalg.norm(w0)
        T = ga * np.array([0.0, 1.0]) + 1j * k * (x0 - focal)
        want = 0
        for mu, c, al in zip(LOBES["mixtureMu"], LOBES["mixtureC"], LOBES["mixtureAlpha"]):
            want += math.exp(_lognorm2(ga) + c + math.log(al)) * _circle_integral(mu * w0 + T)
        got = O.gauss_f(cfg, view, 0, np.append(x0, 0.0), np.append(p2, 0.0), dims=2)
        assert abs(got - want) < 1e-9 * abs(want), (got, want)


def test_fs_zero_lobe_factorises():
    lobes = {"mixtureMu": [0.0], "mixtureC": [-0.7], "mixtureAlpha": [1.0]}
    cfg = _cfg(lobes)
    ga, k = 1 / 0.4 ** 2, 2 * math.pi
    fa, fb = np.array([0.0, 8.0]), np.array([0.5, 8.0])
    for x0 in (np.array([0.1, 8.3]), np.array([0.3, 7.8])):
        aT = ga * np.array([0.0, 1.0]) + 1j * k * (x0 - fa)
        bT = ga * np.array([0.0, 1.0]) + 1j * k * (fb - x0)
        want = math.exp(2 * _lognorm2(ga) - 0.7) * _circle_integral(aT) * _circle_integral(bT)
        got = O.gauss_fs(cfg, 0, 0, np.append(x0, 0.0), dims=2)
        assert abs(got - want) < 1e-9 * abs(want), (got, want)


def test_fs_lobes_approximate_the_double_integral():
    """For gamma != 0 the reference's closed form approximates the double integral
    int int exp(aT.w1 + gamma w1.w2 + bT.w2) dtheta1 dtheta2 (inner integral 2 pi I_0 exactly)."""
    lobes = {"mixtureMu": [4.0], "mixtureC": [-3.0], "mixtureAlpha": [1.0]}
    cfg = _cfg(lobes)
    ga, k = 1 / 0.4 ** 2, 2 * math.pi
    fa, fb = np.array([0.0, 8.0]), np.array([0.5, 8.0])
    x0 = np.array([0.2, 8.1])
    aT = ga * np.array([0.0, 1.0]) + 1j * k * (x0 - fa)
    bT = ga * np.array([0.0, 1.0]) + 1j * k * (fb - x0)
    n = 4096
    th = np.arange(n) * 2 * np.pi / n
    w = np.stack([np.cos(th), np.sin(th)], axis=1)
    inner = 4.0 * w + bT                                      # gamma w1 + bT for every w1
    s = np.sqrt(np.sum(inner * inner, axis=1))
    dbl = np.sum(np.exp(w @ aT) * 2 * np.pi * special.iv(0, s)) * 2 * np.pi / n
    want = math.exp(2 * _lognorm2(ga) - 3.0) * dbl
    got = O.gauss_fs(cfg, 0, 0, np.append(x0, 0.0), dims=2)
    rel = abs(got - want) / abs(want)
    print("2D fs lobe approximation: rel diff %.3f" % rel)
    assert rel < 0.03


def test_first_point_drawn_as_from_point_sources_at_the_focal_points():
    focal = ((-2.0, 3.0), (1.5, 9.0), (0.0, 15.0))
    cfg = _cfg(focal=focal, view=((0.5, 8.0),), batch=256)
    pt = dict(cfg)
    pt["illumination"] = {"type": "PointSource",
                          "location": {"x": np.array([f[0] for f in focal]), "y": np.array([f[1] for f in focal])}}
    pt["view"] = {"type": "PointSource", "location": {"x": np.array([0.5]), "y": np.array([8.0])}}
    a = O.paths(cfg, 2, 0, 256, terms=False)
    b = O.paths(pt, 2, 0, 256, terms=False)
    assert np.array_equal(a["k"], b["k"]) and np.array_equal(a["misses"], b["misses"])
    assert np.array_equal(a["pos"], b["pos"])
    assert a["k"].mean() > 1.5
