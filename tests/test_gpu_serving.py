"""Serving loop: continuous batching keeps every request's result identical to the request
traced alone; admission is by HBM footprint. The mix holds every config family whose arrays the
server must copy (tabular and vMF phase functions, GaussianBeam directions, the correlation
sampler's second source pair) and a request of several renderings."""
import time

import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer, render
from temporalspecklescode_amd.serving import Server
from temporalspecklescode_amd import _native
from temporalspecklescode_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mix():
    a = G.c1_small(4096); a["simulation"]["batchSize"] = 1024; a["simulation"]["fullIteration"] = 4
    b = G.hetero3_small(2048); b["simulation"]["batchSize"] = 512; b["simulation"]["fullIteration"] = 4
    c = G.hetero3_small(1024, point=True); c["simulation"]["batchSize"] = 512; c["simulation"]["fullIteration"] = 2
    d = G.slab_frames(1024, output="correlation"); d["simulation"]["batchSize"] = 256; d["simulation"]["fullIteration"] = 4
    tab = G.with_tabular(G.hetero2_points(256), 2); tab["simulation"]["fullIteration"] = 2
    vmf = G.with_vmf(G.hetero3_small(256)); vmf["simulation"]["fullIteration"] = 2
    beam = W.confocal(focal_x=np.arange(-4.0, 5.0, 2.0), focal_z=np.array([4.0, 12.0]), tau_x=(0.0, 1.0),
                      batch=128, full_iteration=2, seed=3)
    corr = G.corr_sampler(G.hetero3_small(256, point=True), [0.0, 1.0, 2.0], [0.0, 1e-3, 2e-3, 0.0],
                          il2=G.perturbed(G.hetero3_small(256, point=True)["illumination"], 5, 0.5))
    corr["simulation"]["fullIteration"] = 2
    return [(a, 2), (b, 3), (c, 3), (d, 3), (tab, 2), (vmf, 3), (beam, 3), (corr, 3)]


def test_results_identical_alone_and_in_the_mix():
    reqs = _mix()
    alone = []
    for i, (cfg, dims) in enumerate(reqs):
        print("alone", i, flush=True)
        with Tracer(cfg, dims) as tr:
            alone.append(tr.trace())
    with Server(max_inflight=4) as srv:
        ids = []
        for i, (cfg, dims) in enumerate(reqs):
            ids.append(srv.submit(cfg, dims))
            if i in (1, 4):
                time.sleep(0.01)   # later requests arrive while earlier ones are on the device
        outs = []
        for rid in ids:
            srv.wait(rid)
            outs.append(srv.fetch(rid))
        st = srv.stats()
    for i, ((ua, ia), (um, im)) in enumerate(zip(alone, outs)):
        print("request", i, "iter", ia, im, flush=True)
        assert ia == im
        assert np.array_equal(ua, um)
    assert st["completed"] == len(reqs)
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["paths_per_s"] > 0


def test_renderings_match_render():
    """renderingsNum > 1: rendering i at seed + i, laid out like sstmc_render."""
    cfg = G.hetero3_small(512)
    cfg["simulation"]["batchSize"] = 256
    cfg["simulation"]["fullIteration"] = 2
    cfg["simulation"]["renderingsNum"] = 3
    u_ref, it_ref = render(cfg, 3)
    with Server(max_inflight=2) as srv:
        rid = srv.submit(cfg, 3)
        srv.wait(rid)
        u, it = srv.fetch(rid)
    assert u.shape == u_ref.shape and u.shape[-1] == 3
    assert np.array_equal(it, it_ref)
    assert np.array_equal(u, u_ref)
    assert not np.array_equal(u[..., 0], u[..., 1])


def test_admission_by_footprint():
    cfg, dims = _mix()[0]
    with Server(hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(_native.SstmcError) as e:
            srv.submit(cfg, dims)
        assert e.value.code == 9


def test_preloaded_and_cached_media_give_the_same_results():
    """Requests on a preloaded medium, and on a medium another request built, share the server's
    device copy (their configs' grids are not read again) and still give the request's own result."""
    base = G.hetero3_small(512)
    base["simulation"]["batchSize"] = 256
    base["simulation"]["fullIteration"] = 2
    other = G.hetero3_small(512, point=True, seed=8)
    reqs = []
    for i in range(3):
        c = dict(base)
        c["simulation"] = dict(base["simulation"], seed=100 + i)
        reqs.append((c, 3))
    reqs.append((other, 3))
    reqs.append((dict(other, simulation=dict(other["simulation"], seed=77)), 3))
    alone = []
    for cfg, dims in reqs:
        with Tracer(cfg, dims) as tr:
            alone.append(tr.trace())
    with Server(max_inflight=3) as srv:
        srv.preload(base, 3)
        ids = [srv.submit(cfg, dims) for cfg, dims in reqs]
        outs = []
        for rid in ids:
            srv.wait(rid)
            outs.append(srv.fetch(rid))
    for (ua, ia), (um, im) in zip(alone, outs):
        assert ia == im
        assert np.array_equal(ua, um)


def test_mixed_stream_subset_alone_and_served():
    """A subset of the benchmark's mixed request stream (2D and 3D, the three media, L and V in
    [16, 512], 1-8 lags of linear motion, the measured precision form) served with preloaded
    media: every result equal to the request traced alone."""
    stream = W.serving_stream(6, max_paths_log2=16)
    for _, _, cfg in stream:
        cfg["precision"] = {"terms": "fp32", "contraction": "bf16x3", "tauCut": 24.0}
    alone = []
    for i, (_, dims, cfg) in enumerate(stream):
        print("request", i, "dims", dims, "L x V", len(cfg["illumination"]["direction"]["x"]), flush=True)
        with Tracer(cfg, dims) as tr:
            alone.append(tr.trace())
    with Server(max_inflight=4) as srv:
        seen = set()
        for _, dims, cfg in stream:
            key = (dims, len(cfg["medium"]["xAxis"]))
            if key not in seen:
                seen.add(key)
                srv.preload(cfg, dims)
        ids = [srv.submit(cfg, dims) for _, dims, cfg in stream]
        outs = []
        for rid in ids:
            srv.wait(rid)
            outs.append(srv.fetch(rid))
        st = srv.stats()
    for (ua, ia), (um, im) in zip(alone, outs):
        assert ia == im
        assert np.array_equal(ua, um)
    print("p50 %.1f ms p99 %.1f ms" % (st["p50_ms"], st["p99_ms"]), flush=True)
