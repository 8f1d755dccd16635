"""Absolute normalisation of the 3D far-field estimator, pinned by the single-scattering integral.

Per attempt (a path slot, misses included, which is what `iter` counts), the first vertex x is a
scattering event of density q(x) = sigma_s(x) mean_l T_l(x) / A_l (source l uniform, entry point
uniform over its projected rectangle of area A_l, first collision at the free-path density,
absorbed collisions and rays that miss the box are misses), and the path's single-scattering term
is S(l, v) = g_l g_v f(l -> v) e^{i 2 pi rho} / sqrt(p1(x)) with |g_l g_v f|^2 = T_l T_v |f|^2 and
p1 = mean_l T_l / A_l (pathProbabilitySingle, NextEventEstimation.cuh:853-884; FarField.cuh:
469-499). Hence, for every (l, v),

    E[ sum_p |S_p(l, v)|^2 ] / iter = integral over the box of sigma_s T_l(x) T_v(x) |f(d_l . d_v)|^2 dV,

with T_l the transmittance from x back to the box along -d_l and T_v from x to the box along d_v
(HG: |f|^2 = pdf). The right side is evaluated here by quadrature with exact ray-box exit
distances; the left side from the oracle's per-path terms. A wrong projected area, a dropped
sigma, a square root too many or too few in the path weight, or the wrong exit direction of a
transmittance each move the ratio far outside the statistical band.
"""
import math

import numpy as np

import oracle_lib as O

W, H = 12.0, 2.0                      # box [-W, W]^2 x [0, H]
SIGS, SIGA, G = 0.06, 0.02, 0.5


def _dirs(angles):
    return np.array([[math.sin(a) * math.cos(b), math.sin(a) * math.sin(b), math.cos(a)] for a, b in angles])


IL = _dirs([(0.0, 0.0), (0.15, 0.7)])
VW = _dirs([(0.0, 0.0), (0.2, 2.0), (0.1, -1.0)])


def _cfg(batch):
    return {
        "simulation": {"batchSize": batch, "seed": 9, "lambda": 1.0},
        "medium": {"type": "Heterogeneous", "xAxis": np.array([-W, W]), "yAxis": np.array([-W, W]),
                   "zAxis": np.array([0.0, H]), "materialGrid": np.ones((1, 1, 1), np.uint32),
                   "sigs": [0.0, SIGS], "siga": [0.0, SIGA]},
        "scattering": {"type": ["HG", "HG"], "amplitudeFunction": [G, G]},
        "illumination": {"type": "FarField", "direction": {"x": IL[:, 0], "y": IL[:, 1], "z": IL[:, 2]}},
        "view": {"type": "FarField", "direction": {"x": VW[:, 0], "y": VW[:, 1], "z": VW[:, 2]}},
        "sampler": {"type": "TemporalSampler", "t": np.zeros(1), "D": np.zeros(2),
                    "U": {"x": np.zeros(2), "y": np.zeros(2), "z": np.zeros(2)}},
        "tracer": {"type": "nee"},
    }


def _exit(p, d):
    """Distance from points p (n, 3) inside the box to its boundary along d."""
    lo, hi = np.array([-W, -W, 0.0]), np.array([W, W, H])
    t = np.full(p.shape[0], np.inf)
    for a in range(3):
        if d[a] > 0:
            t = np.minimum(t, (hi[a] - p[:, a]) / d[a])
        elif d[a] < 0:
            t = np.minimum(t, (lo[a] - p[:, a]) / d[a])
    return t


def _integral():
    """sigma_s int T_l T_v pdf(d_l . d_v) dV by Gauss-Legendre in z and a fine midpoint grid in x, y."""
    zg, zw = np.polynomial.legendre.leggauss(24)
    z = 0.5 * H * (zg + 1)
    zw = 0.5 * H * zw
    n = 480
    xs = (np.arange(n) + 0.5) / n * 2 * W - W
    X, Y, Z = np.meshgrid(xs, xs, z, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    w = np.broadcast_to((2 * W / n) ** 2 * zw, X.shape).ravel()
    st = SIGS + SIGA
    out = np.zeros((len(IL), len(VW)))
    for l, dl in enumerate(IL):
        Tl = np.exp(-st * _exit(P, -dl))
        for v, dv in enumerate(VW):
            Tv = np.exp(-st * _exit(P, dv))
            c = float(dl @ dv)
            pdf = (1 - G * G) / (4 * math.pi * (1 + G * G - 2 * G * c) ** 1.5)
            out[l, v] = SIGS * pdf * np.sum(w * Tl * Tv)
    return out


def test_single_scattering_intensity_is_the_first_order_integral():
    want = _integral()
    # closed form for the normal pair: T_l T_v = exp(-sigma_t H) everywhere in the box
    pdf1 = (1 - G * G) / (4 * math.pi * (1 - G) ** 3)
    assert abs(want[0, 0] / (SIGS * math.exp(-(SIGS + SIGA) * H) * pdf1 * (2 * W) ** 2 * H) - 1) < 1e-9
    n, chunks = 20000, 8
    means = []
    for c in range(chunks):
        o = O.paths(_cfg(n), 3, c * n, n)
        attempts = n + int(o["misses"].sum())
        s2 = np.abs(o["s"][:, 0]) ** 2                # [p][v][l]
        means.append(s2.sum(axis=0).T / attempts)     # [l][v]
    means = np.array(means)
    got = means.mean(axis=0)
    se = means.std(axis=0, ddof=1) / math.sqrt(chunks)
    z = (got - want) / se
    print("ratio", got / want, "z", z)
    assert np.all(np.abs(z) < 4.5), (got / want, z)
    assert np.all(np.abs(got / want - 1) < 0.02)
