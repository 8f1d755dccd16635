"""Small parity configurations (the oracle finishes each in seconds)."""
import numpy as np

from temporalspecklescode_amd import workloads as W


def c1_small(paths=2048):
    return W.c1(paths=paths, batch=paths)


def hetero3_small(paths=1024, n=16, nl=8, nv=8, g=0.9, point=False, seed=3):
    cfg = W.c2(paths=paths, batch=paths, seed=seed, n=n)
    if not point:
        dirs = W._dirs_grid(int(np.sqrt(nl)) if int(np.sqrt(nl)) ** 2 == nl else 2, 8.0)
        k = len(dirs["x"])
        cfg["illumination"] = {"type": "FarField", "direction": dirs}
        cfg["view"] = {"type": "FarField", "direction": dirs}
    else:
        xy = np.linspace(-n / 2 + 2, n / 2 - 2, 2)
        gx, gy = np.meshgrid(xy, xy, indexing="ij")
        cfg["illumination"] = {"type": "PointSource", "location": {"x": gx.ravel(), "y": gy.ravel(), "z": np.full(4, -n / 2 - 1.0)}}
        cfg["view"] = {"type": "PointSource", "location": {"x": gx.ravel(), "y": gy.ravel(), "z": np.full(4, n / 2 + 1.0)}}
    cfg["scattering"]["amplitudeFunction"] = [g] * 4
    return cfg


def slab_frames(paths=1024, nl=6, nv=5, nt=5, output="field", U=0.0, seed=4):
    cfg = W.c3(paths=paths, batch=paths, seed=seed, nlv=8, nt=nt)
    cfg["illumination"]["direction"] = W._dirs_random(nl, 5.0, seed, 11)
    cfg["view"]["direction"] = W._dirs_random(nv, 5.0, seed, 12)
    cfg["medium"]["zAxis"] = [0.0, 12.0]
    cfg["sampler"]["D"] = np.array([0.0, 2e-3])
    cfg["sampler"]["U"] = {"x": np.array([0.0, U]), "y": np.zeros(2), "z": np.zeros(2)}
    cfg["tracer"] = {"type": "nee", "output": output}
    return cfg


def hetero2_points(paths=256, n=12, seed=5, area=False):
    """2D heterogeneous grid with non-uniform edges and point sources outside the box."""
    edges = np.concatenate([[-6.0], np.sort(np.random.default_rng(seed).uniform(-5.5, 5.5, n - 1)), [6.0]])
    grid = W.random_materials((n, n), seed, 3)
    return {
        "simulation": {"batchSize": paths, "seed": seed, "lambda": 0.7},
        "medium": {"type": "Heterogeneous", "xAxis": edges, "yAxis": edges, "materialGrid": grid,
                   "sigs": [0.0, 0.4, 0.9, 1.6], "siga": [0.0, 0.01, 0.03, 0.02]},
        "scattering": {"type": ["HG"] * 4, "amplitudeFunction": [0.7, 0.7, 0.5, 0.3]},
        "illumination": {"type": "PointSource", "location": {"x": [-2.0, 1.0, 3.0], "y": [-7.0, -7.0, -6.5]},
                         "isAreaSource": area},
        "view": {"type": "PointSource", "location": {"x": [-3.0, 0.0, 2.5, 4.0], "y": [7.0, 6.5, 7.0, 7.5]},
                 "isAreaSource": area},
        "sampler": {"type": "TemporalSampler", "t": [0.0], "D": np.zeros(4),
                    "U": {"x": np.zeros(4), "y": np.zeros(4)}},
        "tracer": {"type": "nee"},
    }


def moving_ff_points(paths=256, seed=6, temporal_amp=True):
    """3D grid, far-field illumination and point views, linear motion + Brownian frames."""
    cfg = hetero3_small(paths, n=12, seed=seed)
    cfg["view"] = {"type": "PointSource", "location": {"x": [-3.0, 0.0, 3.0], "y": [1.0, -2.0, 0.5],
                                                        "z": [7.0, 7.5, 8.0]}}
    cfg["sampler"] = {"type": "TemporalSampler", "t": np.array([0.0, 1.0, 2.5]), "D": np.array([0.0, 1e-3, 0.0, 4e-3]),
                      "U": {"x": np.array([0.0, 0.02, -0.01, 0.0]), "y": np.zeros(4), "z": np.array([0.0, 0.0, 0.03, 0.01])},
                      "isTemporalAmplitude": temporal_amp}
    return cfg


def cbs_backscatter(paths=256, n=12, nd=6, g=0.8, seed=9, point=False):
    """Coherent backscattering, reflection geometry: plane waves entering through z = -n/2 and
    views that are their exact negations (view v = -illumination v), heterogeneous grid; or point
    sources and sensors on the same side of the box."""
    cfg = W.c2(paths=paths, batch=paths, seed=seed, n=n)
    if not point:
        d = W._dirs_random(nd, 10.0, seed, 21)
        cfg["illumination"] = {"type": "FarField", "direction": d}
        cfg["view"] = {"type": "FarField", "direction": {k: -np.asarray(v) for k, v in d.items()}}
    else:
        xy = np.linspace(-n / 2 + 2, n / 2 - 2, 2)
        gx, gy = np.meshgrid(xy, xy, indexing="ij")
        loc = {"x": gx.ravel(), "y": gy.ravel(), "z": np.full(4, -n / 2 - 1.0)}
        cfg["illumination"] = {"type": "PointSource", "location": loc}
        cfg["view"] = {"type": "PointSource", "location": {k: v + 0.5 for k, v in loc.items()}}
    cfg["scattering"]["amplitudeFunction"] = [g] * 4
    cfg["tracer"] = {"type": "nee", "isCBS": True}
    return cfg


def smooth_table(n, dims, k=0):
    """A smooth complex amplitude over the scattering angle ([0, pi] in 3D, [0, 2 pi] in 2D),
    forward-peaked, with a k-dependent shape and phase (k picks one of several materials)."""
    th = np.arange(n) / (n - 1) * (np.pi if dims == 3 else 2 * np.pi)
    mag = 1.0 + 0.6 * np.cos(th) + 0.15 * (k + 1) * np.cos(2 * th) + 1.2 * np.exp(-(th / 0.4) ** 2)
    if dims == 2:
        mag = mag + 1.2 * np.exp(-((th - 2 * np.pi) / 0.4) ** 2)
    return mag * np.exp(1j * (0.5 * (k + 1) * np.sin(th) + 0.3 * k))


def with_tabular(cfg, dims, uniform=False, mixed=False):
    """The same configuration with tabular scattering in every material (one shared table when
    uniform, else one per material); mixed keeps material 2's original phase function."""
    sc = cfg["scattering"]
    types = sc["type"] if isinstance(sc["type"], list) else [sc["type"]]
    nm = len(types)
    amps = list(sc.get("amplitudeFunction", [0.0] * nm))
    new_t, new_a = [], []
    for m in range(nm):
        if mixed and m == 2:
            new_t.append(types[m])
            new_a.append(amps[m])
        else:
            new_t.append("Tabular")
            new_a.append(smooth_table(601 + (0 if uniform else 37 * m), dims, 0 if uniform else m))
    cfg["scattering"] = {"type": new_t, "amplitudeFunction": new_a}
    return cfg


def tabular_mudiff_small(paths=256, nview=24, cbs=True, seed=13):
    """The mudiffCBSValidation2D.m geometry (2D single-cell slab, one plane wave, views around the
    circle) with a smooth synthetic table and fewer views."""
    deg = np.linspace(0, 2 * np.pi, nview)
    cfg = W.mudiff_cbs(smooth_table(721, 2), deg, [-20.0, 20.0], [-20.0, 20.0], 10.0, batch=paths, cbs=cbs,
                       seed=seed)
    return cfg


def _vmf_lobe_c(k):
    """log of the factor that makes one lobe's |exp(k c + C)|^2 integrate to 1 over the sphere."""
    x = 2.0 * k
    return 0.5 * (-np.log(4 * np.pi) if x < 1e-12 else np.log(x) - np.log(4 * np.pi * np.sinh(x)))


def _vmf_mixture(mu, alpha):
    return {"mixtureMu": list(mu), "mixtureC": [_vmf_lobe_c(k) for k in mu], "mixtureAlpha": list(alpha)}


VMF_MIXTURES = [
    _vmf_mixture([1.5, 9.0], [0.6, 0.4]),
    _vmf_mixture([0.0, 3.0, 12.0], [1.0, 0.5, 2.0]),
    _vmf_mixture([4.0], [1.0]),
    _vmf_mixture([2.0, 6.0], [0.3, 0.7]),
]


def with_vmf(cfg, uniform=False, mixed=False):
    """The same configuration with von Mises-Fisher mixtures in every material (one shared mixture
    when uniform, else one per material); mixed keeps material 2's original phase function."""
    sc = cfg["scattering"]
    types = sc["type"] if isinstance(sc["type"], list) else [sc["type"]]
    nm = len(types)
    amps = list(sc.get("amplitudeFunction", [0.0] * nm))
    new_t, new_a = [], []
    for m in range(nm):
        if mixed and m == 2:
            new_t.append(types[m])
            new_a.append(amps[m])
        else:
            new_t.append("vMF")
            new_a.append(VMF_MIXTURES[0 if uniform else m % len(VMF_MIXTURES)])
    cfg["scattering"] = {"type": new_t, "amplitudeFunction": new_a}
    return cfg


def corr_sampler(cfg, lags, D, U=None, il2=None, vw2=None):
    """The same medium, scattering and sources under the TemporalCorrelationSampler (lags `lags`,
    Brownian D and velocity U per material); illumination2 / view2 default to the u1 sources."""
    c = dict(cfg)
    nm = len(D)
    U = np.zeros((nm, 3)) if U is None else np.asarray(U, np.float64)
    c["sampler"] = {"type": "TemporalCorrelationSampler", "t": np.asarray(lags, np.float64),
                    "D": np.asarray(D, np.float64), "U": {"x": U[:, 0], "y": U[:, 1], "z": U[:, 2]}}
    c["tracer"] = {"type": "nee"}
    c["illumination2"] = il2 if il2 is not None else cfg["illumination"]
    c["view2"] = vw2 if vw2 is not None else cfg["view"]
    return c


def perturbed(src, seed, amount):
    """A second source set near `src`: far-field directions tilted by about `amount` radians,
    point locations moved by about `amount`."""
    rng = np.random.default_rng(seed)
    key = "direction" if src["type"] == "FarField" else "location"
    loc = {k: np.array(v, np.float64) for k, v in src[key].items()}
    for k in loc:
        loc[k] = loc[k] + rng.uniform(-amount, amount, loc[k].shape)
    if key == "direction":
        n = np.sqrt(sum(loc[k] ** 2 for k in loc))
        loc = {k: loc[k] / n for k in loc}
    out = dict(src)
    out[key] = loc
    return out
