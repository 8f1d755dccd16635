"""Edge cases of the HIP path against the FP64 CPU restatement on the GPU's own vertices: one
illumination and one view, ragged batch sizes (not a multiple of a wavefront or block), a single
path, source counts that leave partial lane groups and view blocks (65 x 200), and a strongly
absorbing medium where most paths stop at their first vertex (zero multiple-scattering rows)."""
import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer
from test_gpu_parity import _oracle_u, _rel

pytestmark = pytest.mark.gpu


def _dirs(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3)) * [0.12, 0.12, 0.0] + [0.0, 0.0, 1.0]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return {"x": d[:, 0], "y": d[:, 1], "z": d[:, 2]}


def _case(name):
    if name == "one_by_one_ragged":
        cfg = G.hetero3_small(37)
        cfg["illumination"] = {"type": "FarField", "direction": {"x": [0.05], "y": [0.0], "z": [1.0]}}
        cfg["view"] = {"type": "FarField", "direction": {"x": [0.0], "y": [-0.03], "z": [1.0]}}
        return cfg, 3
    if name == "single_path_points_2d":
        cfg = G.hetero2_points(1)
        return cfg, 2
    if name == "partial_groups_65x200":
        cfg = G.hetero3_small(70)
        cfg["illumination"] = {"type": "FarField", "direction": _dirs(65, 1)}
        cfg["view"] = {"type": "FarField", "direction": _dirs(200, 2)}
        return cfg, 3
    if name == "absorbing":
        cfg = G.hetero3_small(300, point=True)
        cfg["medium"]["siga"] = np.array([0.0, 3.0, 3.0, 3.0, 3.0])[:len(cfg["medium"]["sigs"])]
        return cfg, 3
    raise KeyError(name)


@pytest.mark.parametrize("name", ["one_by_one_ragged", "single_path_points_2d", "partial_groups_65x200", "absorbing"])
@pytest.mark.parametrize("terms", ["fp64", "fp32"])
def test_edge_case_matches_oracle(name, terms):
    cfg, dims = _case(name)
    cfg["precision"] = {"terms": terms, "contraction": terms}
    n = cfg["simulation"]["batchSize"]
    with Tracer(cfg, dims) as tr:
        d = tr.debug_batch(0, n)
    k = d["k"]
    print("case", name, terms, "paths", n, "vertices", int(d["off"][-1]), "single-vertex paths",
          int(np.sum(k == 1)), flush=True)
    A, B, S, U = _oracle_u(cfg, dims, d, d["a"], d["b"])
    tol = 1e-10 if terms == "fp64" else 1e-4
    assert _rel(d["u"], U) < tol
    if name == "absorbing":
        assert np.mean(k == 1) > 0.5
        assert np.all(d["b"][k == 1] == 0)
