"""Synthetic workloads (config dicts in the reference's MATLAB layout).

c1..c4 are the four single-request configurations the tracer is measured on; ``serving_stream``
draws the mixed request stream; ``mcx_validation`` rebuilds the configuration of the reference's
``examples/mcxValidation3D.m`` (its ground truth is a committed fixture under tests/golden).
All randomness is counter-based (Philox4x32-10 keyed by a fixed seed), so every generator is
reproducible bit for bit on any machine.
"""
from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------- Philox4x32-10 (vectorised)
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)


def philox4x32(ctr: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """Philox4x32-10 (Salmon et al., SC'11) on an (n, 4) uint32 counter array."""
    c = np.array(ctr, dtype=np.uint32, copy=True).reshape(-1, 4)
    k0, k1 = np.uint32(key[0] & 0xFFFFFFFF), np.uint32(key[1] & 0xFFFFFFFF)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = _M0 * c[:, 0].astype(np.uint64)
        p1 = _M1 * c[:, 2].astype(np.uint64)
        hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & mask).astype(np.uint32)
        hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & mask).astype(np.uint32)
        c = np.stack([hi1 ^ c[:, 1] ^ k0, lo1, hi0 ^ c[:, 3] ^ k1, lo0], axis=1)
        k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return c


def _uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    idx = np.arange(n, dtype=np.uint64)
    ctr = np.zeros((n, 4), np.uint32)
    ctr[:, 0] = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    ctr[:, 1] = (idx >> np.uint64(32)).astype(np.uint32)
    ctr[:, 2] = stream
    w = philox4x32(ctr, (seed & 0xFFFFFFFF, seed >> 32))[:, 0]
    return ((w >> np.uint32(8)).astype(np.float64) + 0.5) / 16777216.0


# ---------------------------------------------------------------- improved Perlin noise
def _perlin_perm(seed: int) -> np.ndarray:
    u = _uniform(seed, 256, stream=0x5045524C)
    perm = np.argsort(u, kind="stable").astype(np.int64)
    return np.concatenate([perm, perm])


def _grad(h, x, y, z):
    # Perlin 2002 gradient selection: 12 cube-edge directions via the low 4 bits of the hash.
    h = h & 15
    u = np.where(h < 8, x, y)
    v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
    return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)


def perlin3(x, y, z, perm):
    """Improved Perlin noise (K. Perlin, "Improving Noise", SIGGRAPH 2002)."""
    xi, yi, zi = (np.floor(x).astype(np.int64) & 255, np.floor(y).astype(np.int64) & 255,
                  np.floor(z).astype(np.int64) & 255)
    x, y, z = x - np.floor(x), y - np.floor(y), z - np.floor(z)
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
    u, v, w = fade(x), fade(y), fade(z)
    A = perm[xi] + yi
    AA, AB = perm[A] + zi, perm[A + 1] + zi
    B = perm[xi + 1] + yi
    BA, BB = perm[B] + zi, perm[B + 1] + zi
    lerp = lambda t, a, b: a + t * (b - a)
    return lerp(w,
                lerp(v, lerp(u, _grad(perm[AA], x, y, z), _grad(perm[BA], x - 1, y, z)),
                     lerp(u, _grad(perm[AB], x, y - 1, z), _grad(perm[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, _grad(perm[AA + 1], x, y, z - 1), _grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u, _grad(perm[AB + 1], x, y - 1, z - 1), _grad(perm[BB + 1], x - 1, y - 1, z - 1))))


def perlin_materials(n: int, seed: int, nmat: int = 4, base_freq: float = 1.0 / 16, octaves: int = 3):
    """(n, n, n) material ids 1..nmat from a 3-octave Perlin field cut at its quantiles."""
    perm = _perlin_perm(seed)
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    f = np.zeros_like(X)
    for o in range(octaves):
        fr = base_freq * (2 ** o)
        f += perlin3(X * fr, Y * fr, Z * fr, perm) * (0.5 ** o)
    qs = np.quantile(f, [(i + 1) / nmat for i in range(nmat - 1)])
    return (1 + np.searchsorted(qs, f, side="right")).astype(np.uint32)


def random_materials(shape, seed: int, nmat: int) -> np.ndarray:
    """i.i.d. uniform material ids 1..nmat from Philox(seed, voxel index), voxel index x fastest."""
    n = int(np.prod(shape))
    u = _uniform(seed, n)
    ids = 1 + np.minimum((u * nmat).astype(np.int64), nmat - 1)
    return ids.reshape(shape, order="F").astype(np.uint32)


def _dirs_grid(n_side: int, max_deg: float):
    th = np.deg2rad(np.linspace(-max_deg, max_deg, n_side))
    tx, ty = np.meshgrid(th, th, indexing="ij")
    d = np.stack([np.tan(tx).ravel(order="F"), np.tan(ty).ravel(order="F"), np.ones(n_side * n_side)])
    d /= np.linalg.norm(d, axis=0)
    return {"x": d[0], "y": d[1], "z": d[2]}


def _dirs_random(n: int, max_deg: float, seed: int, stream: int):
    u = _uniform(seed, 2 * n, stream=stream)
    tx = np.deg2rad((2 * u[:n] - 1) * max_deg)
    ty = np.deg2rad((2 * u[n:] - 1) * max_deg)
    d = np.stack([np.tan(tx), np.tan(ty), np.ones(n)])
    d /= np.linalg.norm(d, axis=0)
    return {"x": d[0], "y": d[1], "z": d[2]}


def _static_sampler(nmat: int, t=(0.0,)):
    return {"type": "TemporalSampler", "t": np.asarray(t, np.float64), "D": np.zeros(nmat),
            "U": {"x": np.zeros(nmat), "y": np.zeros(nmat), "z": np.zeros(nmat)}}


# ---------------------------------------------------------------- the four configurations
def c1(paths: int = 65536, batch: int = 1024, seed: int = 1) -> dict:
    """2D homogeneous slab, 16 x 16 far-field plane waves, one frame."""
    th = np.deg2rad(np.linspace(-10, 10, 16))
    dirs = {"x": np.sin(th), "y": np.cos(th)}
    return {
        "simulation": {"batchSize": batch, "fullIteration": max(1, paths // batch), "seed": seed,
                       "lambda": 1.0},
        "medium": {"type": "Heterogeneous", "xAxis": [-50.0, 50.0], "yAxis": [0.0, 20.0],
                   "materialGrid": np.ones((1, 1), np.uint32), "sigs": [0.0, 1.0], "siga": [0.0, 0.0]},
        "scattering": {"type": ["HG", "HG"], "amplitudeFunction": [0.8, 0.8]},
        "illumination": {"type": "FarField", "direction": dict(dirs)},
        "view": {"type": "FarField", "direction": dict(dirs)},
        "sampler": _static_sampler(2),
        "tracer": {"type": "nee"},
    }


def c2(paths: int = 1 << 20, batch: int = 1 << 16, seed: int = 3, grid_seed: int = 2, n: int = 64) -> dict:
    """3D heterogeneous 64^3 grid, 64 point sources x 64 point sensors."""
    edges = np.arange(-n // 2, n // 2 + 1, dtype=np.float64)
    xy = np.linspace(-28, 28, 8)
    gx, gy = np.meshgrid(xy, xy, indexing="ij")
    src = {"x": gx.ravel(order="F"), "y": gy.ravel(order="F"), "z": np.full(64, -33.0)}
    dst = {"x": gx.ravel(order="F"), "y": gy.ravel(order="F"), "z": np.full(64, 33.0)}
    return {
        "simulation": {"batchSize": batch, "fullIteration": max(1, paths // batch), "seed": seed,
                       "lambda": 0.5},
        "medium": {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "zAxis": edges,
                   "materialGrid": random_materials((n, n, n), grid_seed, 3),
                   "sigs": [0.0, 0.5, 1.0, 2.0], "siga": [0.0, 0.01, 0.02, 0.05]},
        "scattering": {"type": ["HG"] * 4, "amplitudeFunction": [0.9] * 4},
        "illumination": {"type": "PointSource", "location": src},
        "view": {"type": "PointSource", "location": dst},
        "sampler": _static_sampler(4),
        "tracer": {"type": "nee"},
    }


def c3(paths: int = 1 << 22, batch: int = 1 << 15, seed: int = 4, nlv: int = 128, nt: int = 9) -> dict:
    """3D slab with Brownian scatterers: 128 x 128 plane waves, correlation over 8 lags."""
    D = (0.01 ** 2) / (2 * 3 * 1.0)  # sqrt(2*DIMS*D*dt) = 0.01 lambda for dt = 1
    return {
        "simulation": {"batchSize": batch, "fullIteration": max(1, paths // batch), "seed": seed,
                       "lambda": 1.0},
        "medium": {"type": "Heterogeneous", "xAxis": [-64.0, 64.0], "yAxis": [-64.0, 64.0],
                   "zAxis": [0.0, 40.0], "materialGrid": np.ones((1, 1, 1), np.uint32),
                   "sigs": [0.0, 1.0], "siga": [0.0, 0.001]},
        "scattering": {"type": ["HG", "HG"], "amplitudeFunction": [0.85, 0.85]},
        "illumination": {"type": "FarField", "direction": _dirs_random(nlv, 5.0, seed, 0x494C)},
        "view": {"type": "FarField", "direction": _dirs_random(nlv, 5.0, seed, 0x5657)},
        "sampler": {"type": "TemporalSampler", "t": np.arange(nt, dtype=np.float64),
                    "D": np.array([0.0, D]), "U": {"x": np.zeros(2), "y": np.zeros(2), "z": np.zeros(2)}},
        "tracer": {"type": "nee", "output": "correlation"},
    }


def c3_analytic(paths: int = 1 << 22, batch: int = 1 << 15, seed: int = 4, nlv: int = 128, nt: int = 9) -> dict:
    """The c3 slab under the TemporalCorrelationSampler (analytic decorrelation; illumination2 =
    illumination, view2 = view, lags 0..nt-1)."""
    cfg = c3(paths=paths, batch=batch, seed=seed, nlv=nlv, nt=nt)
    cfg["sampler"] = dict(cfg["sampler"], type="TemporalCorrelationSampler")
    cfg["tracer"] = {"type": "nee"}
    cfg["illumination2"] = cfg["illumination"]
    cfg["view2"] = cfg["view"]
    return cfg


def spatio_temporal_2d(amp, pixels: int = 51, motion: str = "Brownian-Linear", batch: int = 100,
                       full_iteration: int = 100, seed: int = 1, mfp: float = 40.0, ux: float = 1.0,
                       D: float = 0.015) -> dict:
    """The configuration of the reference's examples/spatioTemporalCorrelation2D.m (lines 8-108): a
    2D single-cell slab [-50, 50] x [-20, 20] of mean free path `mfp` with the tabular amplitude
    `amp` in both materials; TemporalCorrelationSampler with the u1 illumination and view fixed at
    the first pixel and the u2 ones swept over `pixels` positions 80 apart, 30 below the slab
    (reflection); lags 0..2; Brownian D and/or linear velocity ux along x by `motion`."""
    x = np.linspace(-1.0, 1.0, pixels)
    sep, dil, dvw = 80.0, 30.0, 30.0

    def pts(xs, y):
        return {"type": "PointSource", "location": {"x": np.asarray(xs, np.float64), "y": np.full(pixels, y)}}

    sig = np.array([0.0, 1.0 / mfp])
    return {
        "simulation": {"batchSize": batch, "fullIteration": full_iteration, "seed": seed, "lambda": 1.0},
        "scattering": {"type": ["Tabular", "Tabular"], "amplitudeFunction": [amp, amp]},
        "medium": {"type": "Heterogeneous", "xAxis": [-50.0, 50.0], "yAxis": [-20.0, 20.0],
                   "materialGrid": np.ones((1, 1), np.uint32), "sigs": sig, "siga": sig * 0},
        "illumination": pts(0 * x + x[0] - sep / 2, -dil),
        "illumination2": pts(x - sep / 2, -dil),
        "view": pts(0 * x + x[0] + sep / 2, -dvw),
        "view2": pts(x + sep / 2, -dvw),
        "sampler": {"type": "TemporalCorrelationSampler", "t": np.linspace(0.0, 2.0, pixels),
                    "D": np.array([0.0, D if motion != "Linear" else 0.0]),
                    "U": {"x": np.array([0.0, ux if motion != "Brownian" else 0.0]), "y": np.zeros(2)}},
        "tracer": {"type": "nee"},
    }


CONFOCAL_LOBES = [   # examples/confocalRenderings.m lines 16-47: vMF lobes of the five materials
    {"mixtureMu": [0.0], "mixtureC": [-1.2655121234846453964889], "mixtureAlpha": [1.0]},
    {"mixtureMu": [0, 7.864623088553429, 1.432699943392017, 34.940513109419960, 1.582028646657162e+02,
                   7.527763768351407e+02, -0.026667703987043],
     "mixtureC": [-2.531024246969291, -7.640125542225978, -2.852369324003370, -33.224743185787034,
                  -1.549768635690816e+02, -7.479904856941710e+02, -2.531142771898976],
     "mixtureAlpha": [0.096076950725656, 0.250010487228936, 0.504158076691359, 0.163528156678287,
                      0.097875866684034, 0.031641230957240, 0.223222884401916]},
    {"mixtureMu": [0, 1.764406405245083e+02, 1.382801893975425, 36.855826580175176, 7.949781817711435,
                   -0.078158767086068],
     "mixtureC": [-2.531024246969291, -1.731055330854194e+02, -2.831561238911424, -35.086689924296930,
                  -7.714514400144169, -2.532042171876453],
     "mixtureAlpha": [0.105617258338313, 0.067057292201984, 0.758751296773679, 0.206483373674102,
                      0.359274660451266, 0.326443781481564]},
    {"mixtureMu": [0, 8.567948260089990, 40.747148947304666, 1.512176451996613, -0.006999141602579],
     "mixtureC": [-2.531024246969291, -8.257797032110451, -38.877640140958086, -2.886694147327419,
                  -2.531032411619821],
     "mixtureAlpha": [0.253871838456329, 0.451553613239657, 0.137923738827578, 1.057954808434358,
                      0.476692395924039]},
    {"mixtureMu": [0, 11.579090874816515, 1.779309977673783, 67.550452715965830, 4.215886314000809e+02,
                   2.882436622668382e+03, -0.005285363127619],
     "mixtureC": [-2.531024246969291, -10.967766980382587, -3.012069941353215, -65.175455016052640,
                  -4.173824784349965e+02, -2.876308108469822e+03, -2.531028902808854],
     "mixtureAlpha": [0.063275130856032, 0.164645985522835, 0.319201646728633, 0.107094124396163,
                      0.065917372206027, 0.026062432689170, 0.164887237242181]},
]


def confocal(focal_x=np.arange(-50.0, 51.0, 1.0), focal_z=np.arange(0.0, 21.0, 2.0), tau_x=(-1.0, 0.0, 1.0),
             batch: int = 50, full_iteration: int = 100, seed: int = 1, aperture: float = 0.4) -> dict:
    """The configuration of the reference's examples/confocalRenderings.m (lines 8-95): four cells of
    vMF-mixture materials in [-100, 100] x [-100, 100] x [0, 20], GaussianBeam illuminations focused
    on an (x, z) grid along +z, GaussianBeam views on the same grid shifted by tau_x in x."""
    fx, fz = np.asarray(focal_x, np.float64), np.asarray(focal_z, np.float64)
    xl, zl = np.meshgrid(fx, fz, indexing="ij")
    xv, zv, tv = np.meshgrid(fx, fz, np.asarray(tau_x, np.float64), indexing="ij")
    xl, zl = xl.ravel(order="F"), zl.ravel(order="F")
    xv, zv, tv = xv.ravel(order="F"), zv.ravel(order="F"), tv.ravel(order="F")
    sigs = np.array([0, 1 / 5, 1 / 10, 1 / 4, 1 / 20])

    def beams(x, z):
        n = x.size
        return {"type": "GaussianBeam", "focalPoint": {"x": x, "y": np.zeros(n), "z": z},
                "focalDirection": {"x": np.zeros(n), "y": np.zeros(n), "z": np.ones(n)}, "aperture": aperture}

    return {
        "simulation": {"batchSize": batch, "fullIteration": full_iteration, "seed": seed, "lambda": 1.0},
        "scattering": {"type": ["vMF"] * 5, "amplitudeFunction": [dict(x) for x in CONFOCAL_LOBES]},
        "medium": {"type": "Heterogeneous", "xAxis": [-100.0, 0.0, 100.0], "yAxis": [-100.0, 100.0],
                   "zAxis": [0.0, 10.0, 20.0], "materialGrid": np.array([1, 2, 3, 4], np.uint32).reshape(2, 1, 2, order="F"),
                   "sigs": sigs, "siga": sigs * 0},
        "illumination": beams(xl, zl),
        "view": beams(xv + tv, zv),
        "sampler": _static_sampler(5),
        "tracer": {"type": "nee"},
    }


def gauss_transmission_matrix(pixels: int = 51, dp: float = 0.2, backscattering: bool = True, batch: int = 1,
                              full_iteration: int = 1000, seed: int = 1, aperture: float = 0.4) -> dict:
    """The configuration of the reference's examples/transmissionMatrixRendering.m (lines 1-95): a
    2000 x 2000 x 100 slab of one material (mean free path 40, isotropic vMF lobes), pixels x pixels
    GaussianBeam illuminations focused on the z = 0 face along +z, GaussianBeam views on the same
    grid (backscattering: at z = 0 looking along -z; transmissive: at z = 100 along +z), five
    frames of Brownian motion."""
    r = (np.arange(pixels, dtype=np.float64) - (pixels - 1) / 2) * dp
    X, Y = np.meshgrid(r, r, indexing="ij")
    x, y = X.ravel(order="F"), Y.ravel(order="F")
    n = x.size
    iso = {"mixtureMu": [0.0], "mixtureC": [-1.2655121234846453964889], "mixtureAlpha": [1.0]}
    sigs = np.array([0.0, 1 / 40])
    vz, vd = (0.0, -1.0) if backscattering else (100.0, 1.0)
    return {
        "simulation": {"batchSize": batch, "fullIteration": full_iteration, "seed": seed, "lambda": 1.0},
        "scattering": {"type": ["vMF"] * 2, "amplitudeFunction": [dict(iso), dict(iso)]},
        "medium": {"type": "Heterogeneous", "xAxis": [-1000.0, 1000.0], "yAxis": [-1000.0, 1000.0],
                   "zAxis": [0.0, 100.0], "materialGrid": np.ones((1, 1, 1), np.uint32), "sigs": sigs,
                   "siga": sigs * 0},
        "illumination": {"type": "GaussianBeam", "focalPoint": {"x": x, "y": y, "z": np.zeros(n)},
                         "focalDirection": {"x": np.zeros(n), "y": np.zeros(n), "z": np.ones(n)},
                         "aperture": aperture},
        "view": {"type": "GaussianBeam", "focalPoint": {"x": x, "y": y, "z": np.full(n, vz)},
                 "focalDirection": {"x": np.zeros(n), "y": np.zeros(n), "z": np.full(n, vd)},
                 "aperture": aperture},
        "sampler": {"type": "TemporalSampler", "t": np.arange(0.0, 1.01, 0.25), "D": np.array([0.0, 1e-3]),
                    "U": {"x": np.zeros(2), "y": np.zeros(2), "z": np.zeros(2)}},
        "tracer": {"type": "nee"},
    }


def c4(paths: int = 1 << 24, batch: int = 1 << 16, seed: int = 6, perlin_seed: int = 5, n: int = 128,
       nside: int = 32) -> dict:
    """Transmission matrix: 128^3 Perlin-thresholded grid, 1024 x 1024 plane waves."""
    edges = np.arange(-n // 2, n // 2 + 1, dtype=np.float64)
    dirs = _dirs_grid(nside, 8.0)
    return {
        "simulation": {"batchSize": batch, "fullIteration": max(1, paths // batch), "seed": seed,
                       "lambda": 1.0},
        "medium": {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "zAxis": edges,
                   "materialGrid": perlin_materials(n, perlin_seed),
                   "sigs": [0.0, 0.2, 1.0, 3.0, 5.0], "siga": [0.0, 0.01, 0.01, 0.02, 0.05]},
        "scattering": {"type": ["HG"] * 5, "amplitudeFunction": [0.9] * 5},
        "illumination": {"type": "FarField", "direction": dict(dirs)},
        "view": {"type": "FarField", "direction": dict(dirs)},
        "sampler": _static_sampler(5),
        "tracer": {"type": "nee"},
    }


def mcx_validation(t, D, MFP, R, h, dims, viSep, batch=10000, full_iteration=1, seed=1) -> dict:
    """The configuration built by the reference's examples/mcxValidation3D.m (lines 13-97)."""
    lam = 500e-9 * 100
    xa = list(np.arange(-dims[0] / 2, dims[0] / 2 + 1e-12, h / 2))
    if xa[-1] != dims[0] / 2:
        xa.append(dims[0] / 2)
    xa = np.array(xa)
    odd = xa[0::2]
    if xa[1::2].size == odd.size:
        xa[1::2] = odd + R
    else:
        xa[1::2] = odd[:-1] + R
    za = list(np.arange(0, dims[2] + 1e-12, h / 2))
    if za[-1] != dims[2]:
        za.append(dims[2])
    za = np.array(za)
    oddz = za[0::2]
    if za[1::2].size == oddz.size:
        za[1::2] = oddz + R
    else:
        za[1::2] = oddz[:-1] + R
    ya = np.array([-dims[1] / 2, dims[1] / 2])
    grid = np.ones((xa.size - 1, 1, za.size - 1), np.uint32)
    grid[0::2, :, 0::2] = 2
    sigs = np.array([0.0, 1.0 / MFP, 1.0 / MFP])
    return {
        "simulation": {"batchSize": batch, "fullIteration": full_iteration, "seed": seed, "lambda": lam},
        "medium": {"type": "Heterogeneous", "xAxis": xa, "yAxis": ya, "zAxis": za, "materialGrid": grid,
                   "sigs": sigs, "siga": sigs * 0},
        "scattering": {"type": ["HG"] * 3, "amplitudeFunction": [0.0, 0.0, 0.0]},
        "illumination": {"type": "PointSource", "location": {"x": [-viSep / 2], "y": [0.0], "z": [0.0]}},
        "view": {"type": "PointSource", "location": {"x": [viSep / 2], "y": [0.0], "z": [0.0]}},
        "sampler": {"type": "TemporalSampler", "t": np.asarray(t, np.float64).ravel(),
                    "D": np.array([0.0, 0.0, D]), "U": {"x": np.zeros(3), "y": np.zeros(3), "z": np.zeros(3)}},
        "tracer": {"type": "nee", "output": "correlation"},
    }


def mudiff_cbs(amp, deg, x, y, mfp, batch: int = 1000, renderings: int = 1, cbs: bool = True, seed: int = 1) -> dict:
    """The field-output configuration of the reference's examples/mudiffCBSValidation2D.m (lines 10-57,
    96-117): a 2D single-cell slab of mean free path `mfp` with the tabular amplitude function
    `amp` in both materials, one far-field illumination along +y and far-field views at the
    angles `deg` (radians) from +y, with or without coherent backscattering."""
    amp = np.asarray(amp, np.complex128).ravel()
    deg = np.asarray(deg, np.float64).ravel()
    sigs = np.array([0.0, 1.0 / float(mfp)])
    return {
        "simulation": {"batchSize": batch, "fullIteration": 1, "renderingsNum": renderings, "seed": seed,
                       "lambda": 1.0},
        "medium": {"type": "Heterogeneous", "xAxis": np.asarray(x, np.float64).ravel(),
                   "yAxis": np.asarray(y, np.float64).ravel(), "materialGrid": np.ones((1, 1), np.uint32),
                   "sigs": sigs, "siga": sigs * 0},
        "scattering": {"type": ["Tabular", "Tabular"], "amplitudeFunction": [amp, amp]},
        "illumination": {"type": "FarField", "direction": {"x": [0.0], "y": [1.0]}},
        "view": {"type": "FarField", "direction": {"x": np.sin(deg), "y": np.cos(deg)}},
        "sampler": _static_sampler(2),
        "tracer": {"type": "nee", "isCBS": bool(cbs)},
    }


def serving_stream(n_requests: int, seed: int = 7, max_paths_log2: int = 22, max_lv: int = 512):
    """Mixed request stream: (arrival_offset_s, dims, config) tuples.

    Each request picks 2D or 3D, one of three preloaded media (32^3, 64^3, 128^3), L and V
    log-uniform in [16, max_lv], paths log-uniform in [2^16, 2^max_paths_log2], 1-8 lags of linear
    motion with velocity uniform in +-0.05 lambda/frame, far-field or point sources.
    """
    u = _uniform(seed, 16 * n_requests, stream=0x53525631).reshape(n_requests, 16)
    media3 = {}
    out = []
    t = 0.0
    for i in range(n_requests):
        r = u[i]
        t += -math.log(r[0]) / 20.0  # Poisson arrivals, 20 requests/s
        dims = 2 if r[1] < 0.25 else 3
        mi = int(r[2] * 3)
        n = (32, 64, 128)[mi]
        L = int(round(math.exp(math.log(16) + r[3] * (math.log(max_lv) - math.log(16)))))
        V = int(round(math.exp(math.log(16) + r[4] * (math.log(max_lv) - math.log(16)))))
        lp = 16 + r[5] * (max_paths_log2 - 16)
        paths = int(2 ** lp)
        lags = 1 + int(r[6] * 8)
        sigs_hi = 0.2 + r[7] * 4.8
        g = 0.7 + r[8] * 0.25
        vel = (2 * r[9] - 1) * 0.05
        batch = min(1 << 14, 1 << int(math.floor(math.log2(paths))))
        nmat = 5
        sigs = np.array([0.0] + list(np.linspace(0.2, sigs_hi, nmat - 1)))
        siga = 0.01 * sigs
        edges = np.arange(-n // 2, n // 2 + 1, dtype=np.float64)
        if dims == 3:
            key = n
            if key not in media3:
                media3[key] = random_materials((n, n, n), 100 + n, nmat - 1)
            grid = media3[key]
            medium = {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "zAxis": edges,
                      "materialGrid": grid, "sigs": sigs, "siga": siga}
            ill = {"type": "FarField", "direction": _dirs_random(L, 8.0, seed + i, 1)}
            if r[10] < 0.5:
                vw = {"type": "FarField", "direction": _dirs_random(V, 8.0, seed + i, 2)}
            else:
                xs = (u[i, 11] * 2 - 1) * (n / 2 - 1) * np.linspace(-1, 1, V)
                vw = {"type": "PointSource", "location": {"x": xs, "y": np.zeros(V), "z": np.full(V, n / 2 + 1.0)}}
        else:
            grid = random_materials((n, n), 200 + n, nmat - 1)
            medium = {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "materialGrid": grid,
                      "sigs": sigs, "siga": siga}
            thl = np.deg2rad(np.linspace(-8, 8, L))
            thv = np.deg2rad(np.linspace(-8, 8, V))
            ill = {"type": "FarField", "direction": {"x": np.sin(thl), "y": np.cos(thl)}}
            vw = {"type": "FarField", "direction": {"x": np.sin(thv), "y": np.cos(thv)}}
        nt = lags + 1
        Ux = np.full(nmat, vel)
        Ux[0] = 0.0
        cfg = {
            "simulation": {"batchSize": batch, "fullIteration": max(1, paths // batch), "seed": 1000 + i,
                           "lambda": 1.0},
            "medium": medium,
            "scattering": {"type": ["HG"] * nmat, "amplitudeFunction": [g] * nmat},
            "illumination": ill,
            "view": vw,
            "sampler": {"type": "TemporalSampler", "t": np.arange(nt, dtype=np.float64), "D": np.zeros(nmat),
                        "U": {"x": Ux, "y": np.zeros(nmat), "z": np.zeros(nmat)}},
            "tracer": {"type": "nee"},
        }
        out.append((t, dims, cfg))
    return out
