"""Host-side checks of the synthetic workload generators (no GPU)."""
import numpy as np

from temporalspecklescode_amd import workloads as W
from temporalspecklescode_amd.config import CConfig, sizes


def test_serving_stream_requests_are_valid_configs():
    stream = W.serving_stream(64, max_paths_log2=18)
    assert len(stream) == 64
    arrivals = [a for a, _, _ in stream]
    assert all(b > a for a, b in zip(arrivals, arrivals[1:]))
    for _, dims, cfg in stream:
        CConfig(cfg, dims)   # raises on any invalid field
        nl, nv, nt, _ = sizes(cfg, dims)
        assert 16 <= nl <= 512 and 16 <= nv <= 512 and 2 <= nt <= 9


def test_random_materials_cover_ids_one_to_n():
    g = W.random_materials((16, 16, 16), 3, 4)
    assert g.min() == 1 and g.max() == 4
    assert np.array_equal(g, W.random_materials((16, 16, 16), 3, 4))
