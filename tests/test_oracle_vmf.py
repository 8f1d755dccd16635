"""von Mises-Fisher mixture phase functions in the CPU restatement (Scattering.cuh:613-884,
895-1110; SpecialFunctions.cuh:187-205; VectorType.cuh:169-205).

Pinned by what the mathematics fixes, independently of the restatement's own tables:
  * one lobe with C = 0 and alpha = 1 has the amplitude exp(kappa c), and its sampled cosine follows
    the textbook vMF law of |f|^2, concentration 2 kappa: F(c) = (e^{2 kappa c} - e^{-2 kappa}) /
    (e^{2 kappa} - e^{-2 kappa});
  * for mixtures the pdf is |f|^2 normalised over the sphere (2 pi int |f|^2 dc, by quadrature),
    and the sampled cosines follow the CDF integrated from that |f|^2 (KS), mean cosine included;
  * in 2D a vMF material scatters isotropically (Scattering.cuh:898-903, 963-968, 1094-1099).
"""
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle_lib as O
from test_oracle_tabular import slab


def vmf_slab(dims, mixtures):
    cfg = slab(dims, np.ones(4, np.complex128))
    cfg["scattering"] = {"type": ["vMF", "vMF"], "amplitudeFunction": [mixtures, mixtures]}
    return cfg


MIXTURES = {
    "one_lobe": {"mixtureMu": [4.0], "mixtureC": [0.0], "mixtureAlpha": [1.0]},
    "two_lobes": {"mixtureMu": [1.5, 9.0], "mixtureC": [0.2, -3.0], "mixtureAlpha": [0.6, 0.4]},
    "three_lobes": {"mixtureMu": [0.0, 3.0, 12.0], "mixtureC": [-1.0, -2.0, -6.0], "mixtureAlpha": [1.0, 0.5, 2.0]},
}


def amp2(cfg, c):
    return np.abs(O.amplitude_complex(cfg, 3, 1, c)) ** 2


def test_one_lobe_amplitude_and_law():
    k = 4.0
    cfg = vmf_slab(3, MIXTURES["one_lobe"])
    c = np.linspace(-1, 1, 21)
    f = O.amplitude_complex(cfg, 3, 1, c)
    assert np.allclose(f.real, np.exp(k * c), rtol=1e-14) and np.all(f.imag == 0)
    prev = np.array([0.2, 0.7, -0.68])
    prev /= np.linalg.norm(prev)
    d = O.sample_directions(cfg, 3, 1, prev, 40000)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12)
    # the law is |f|^2 = exp(2 kappa c): vMF with concentration 2 kappa
    cdf = lambda x: (np.exp(2 * k * x) - math.exp(-2 * k)) / (math.exp(2 * k) - math.exp(-2 * k))
    assert stats.kstest(d @ prev, cdf).pvalue > 1e-3


def large_kappa_error(mix):
    # the reference normalises a lobe of concentration x >= 5 by log(x) - log(2 pi) - x instead of
    # log(x / (4 pi sinh x)) (SpecialFunctions.cuh:198-200): a relative error 1 - exp(-2 x)
    mu = np.asarray(mix["mixtureMu"], float)
    pairs = (mu[:, None] + mu[None, :]).ravel()
    return 2 * float(np.max(np.where(pairs >= 5, np.exp(-2 * pairs), 0.0))) + 1e-10


@pytest.mark.parametrize("name", sorted(MIXTURES))
def test_pdf_is_normalised_amplitude_squared(name):
    cfg = vmf_slab(3, MIXTURES[name])
    tol = large_kappa_error(MIXTURES[name])
    pdf = lambda c: O.phase(cfg, 3, 1, c, pdf=True)
    norm = 2 * math.pi * integrate.quad(pdf, -1, 1, limit=200, epsabs=0, epsrel=1e-12)[0]
    assert abs(norm - 1) < tol
    w = 2 * math.pi * integrate.quad(lambda c: amp2(cfg, [c])[0], -1, 1, limit=200, epsabs=0, epsrel=1e-12)[0]
    c = np.linspace(-1, 1, 41)
    p = np.array([pdf(x) for x in c])
    assert np.allclose(p, amp2(cfg, c) / w, rtol=tol, atol=0)


@pytest.mark.parametrize("name", ["two_lobes", "three_lobes"])
def test_mixture_sampling_follows_amplitude_squared(name):
    cfg = vmf_slab(3, MIXTURES[name])
    grid = np.linspace(-1, 1, 8001)
    dens = amp2(cfg, grid)
    cdfv = np.concatenate([[0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    cdfv /= cdfv[-1]
    prev = np.array([0.0, 0.6, 0.8])
    d = O.sample_directions(cfg, 3, 1, prev, 40000)
    c = d @ prev
    assert stats.kstest(c, lambda x: np.interp(x, grid, cdfv)).pvalue > 1e-3
    mean = integrate.trapezoid(grid * dens, grid) / integrate.trapezoid(dens, grid)
    assert abs(c.mean() - mean) < 5 * c.std() / math.sqrt(c.size)


def test_vmf_is_isotropic_in_2d():
    cfg = vmf_slab(2, MIXTURES["two_lobes"])
    iso_amp, iso_pdf = 1 / math.sqrt(2 * math.pi), 1 / (2 * math.pi)
    for c in (-0.9, 0.0, 0.4, 1.0):
        assert abs(O.phase(cfg, 2, 1, c) - iso_amp) < 1e-15
        assert abs(O.phase(cfg, 2, 1, c, pdf=True) - iso_pdf) < 1e-15
    d = O.sample_directions(cfg, 2, 1, np.array([0.6, 0.8, 0.0]), 20000)
    ang = np.arctan2(d[:, 1], d[:, 0])
    assert stats.kstest(ang, stats.uniform(loc=-math.pi, scale=2 * math.pi).cdf).pvalue > 1e-3
