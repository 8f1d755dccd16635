"""The contraction stage replayed from captured HIP graphs (simulation.graphs): the first batch of
each buffer set runs eagerly and is captured, every later batch replays the graphs; the renderings
must equal the eager ones bit for bit. Several batches per rendering (each set's graphs replayed
several times), two renderings per tracer (replays after the tracer's own eager work and
synchronisation), across the contraction forms: bf16x3 and FP64, field CBS, the correlation sampler with
its CBS pairings, GaussianBeam single scattering."""
import numpy as np
import pytest

import gpu_cases as G
from temporalspecklescode_amd import Tracer
from temporalspecklescode_amd import workloads as W

pytestmark = pytest.mark.gpu


def _cases():
    tm = W.c4(paths=8 * 512, batch=512)
    cbs = G.cbs_backscatter(paths=256)
    acorr = G.corr_sampler(G.cbs_backscatter(256), [0.0, 0.5, 1.5], [0.0, 0.01, 0.02, 0.005])
    acorr["tracer"] = dict(acorr["tracer"], isCBS=True)
    gauss = W.confocal(focal_x=np.arange(-4.0, 5.0, 2.0), focal_z=np.array([4.0, 12.0]), tau_x=(0.0, 1.0),
                       batch=128, full_iteration=1, seed=1)
    return {"tm_bf16x3": (tm, 3, None), "tm_fp64": (tm, 3, {"terms": "fp64", "contraction": "fp64"}),
            "cbs_field": (cbs, 3, None), "corr_cbs": (acorr, 3, {"terms": "fp64", "contraction": "fp64"}),
            "gauss": (gauss, 3, None)}


CASES = _cases()


def _render(cfg, dims, graphs, seeds=(3, 4)):
    c = dict(cfg, simulation=dict(cfg["simulation"], graphs=graphs))
    out = []
    with Tracer(c, dims) as tr:
        for s in seeds:
            u, it = tr.trace(seed=s)
            t = tr.timing()
            out.append((u.copy(), it, t["graph_replays"], t["pipeline_sets"]))
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replays_equal_eager(case):
    cfg, dims, prec = CASES[case]
    cfg = dict(cfg)
    if prec:
        cfg["precision"] = prec
    batch = cfg["simulation"]["batchSize"]
    cfg["simulation"] = dict(cfg["simulation"], fullIteration=6, batchSize=min(batch, 256))
    eager = _render(cfg, dims, False)
    graph = _render(cfg, dims, True)
    for i, ((ue, ie, re, _), (ug, ig, rg, _)) in enumerate(zip(eager, graph)):
        print(case, "norm", np.linalg.norm(ue), "max diff", np.abs(ue - ug).max(), "replays", rg, flush=True)
        assert ie == ig
        assert np.array_equal(ue, ug)
        assert re == 0
        # 6 batches over the pipeline's sets: the first of each set captured in the first rendering
        sets = graph[i][3]
        assert sets in (2, 3)
        assert rg == (6 - sets if i == 0 else 6)
    assert not np.array_equal(eager[0][0], eager[1][0])   # the two seeds differ
