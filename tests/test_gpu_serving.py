"""Serving loop: continuous batching keeps every request's result identical to the request
traced alone; admission is by HBM footprint."""
import time

import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer
from temporalspecklescode_amd.serving import Server
from temporalspecklescode_amd import _native

pytestmark = pytest.mark.gpu


def _mix():
    a = G.c1_small(4096); a["simulation"]["batchSize"] = 1024; a["simulation"]["fullIteration"] = 4
    b = G.hetero3_small(2048); b["simulation"]["batchSize"] = 512; b["simulation"]["fullIteration"] = 4
    c = G.hetero3_small(1024, point=True); c["simulation"]["batchSize"] = 512; c["simulation"]["fullIteration"] = 2
    d = G.slab_frames(1024, output="correlation"); d["simulation"]["batchSize"] = 256; d["simulation"]["fullIteration"] = 4
    return [(a, 2), (b, 3), (c, 3), (d, 3)]


def test_results_identical_alone_and_in_the_mix():
    reqs = _mix()
    alone = []
    for cfg, dims in reqs:
        with Tracer(cfg, dims) as tr:
            alone.append(tr.trace())
    with Server(max_inflight=4) as srv:
        ids = []
        for i, (cfg, dims) in enumerate(reqs):
            ids.append(srv.submit(cfg, dims))
            if i == 1:
                time.sleep(0.01)   # later requests arrive while earlier ones are on the device
        outs = []
        for rid in ids:
            srv.wait(rid)
            outs.append(srv.fetch(rid))
        st = srv.stats()
    for (ua, ia), (um, im) in zip(alone, outs):
        assert ia == im
        assert np.array_equal(ua, um)
    assert st["completed"] == len(reqs)
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["paths_per_s"] > 0


def test_admission_by_footprint():
    cfg, dims = _mix()[0]
    with Server(hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(_native.SstmcError) as e:
            srv.submit(cfg, dims)
        assert e.value.code == 9
