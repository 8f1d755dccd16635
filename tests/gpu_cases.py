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
