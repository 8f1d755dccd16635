"""The reference's own known answer on the GPU path: the MCX temporal-correlation curve of
examples/mcxValidation3D.m (fixture tests/golden/mcx_validation.json), traced with the FP32
transport / FP32 terms / bf16x3 contraction."""
import json
import os

import numpy as np
import pytest

from temporalspecklescode_amd import Tracer, workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mcx_validation.json")


def test_mcx_curve_gpu():
    r = json.load(open(GOLD))
    mcx = np.array(r["mcx_res"])
    cfg = W.mcx_validation(r["t"], r["D"], r["MFP"], r["R"], r["h"], r["dims"], r["viSep"], batch=16384,
                           full_iteration=16, seed=77)
    res = []
    with Tracer(cfg, 3) as tr:
        for s in range(12):
            u, it = tr.trace(seed=77 + s)
            res.append(u[0, 0, :].real / it)
    res = np.array(res)
    m = res.mean(0)
    se = res.std(0, ddof=1) / np.sqrt(len(res))
    z = (m - mcx) / se
    print("C(0) %.5f +- %.5f (mcx %.5f); max |z| %.2f" % (m[0], se[0], mcx[0], np.abs(z).max()), flush=True)
    assert np.all(np.abs(z) < 4), z
