"""The MATLAB front end on the GPU: `python -m temporalspecklescode_amd.matlab cfg.mat --out r.mat`
writes u (Nl x Nv x Nt x renderings) and iter (1 x renderings) equal to sstmc3D on the same
config."""
import os

import numpy as np
import pytest
import scipy.io as sio

import gpu_cases as G
from temporalspecklescode_amd import matlab as M
from temporalspecklescode_amd import sstmc3D
from test_matlab import _mat, _cell

pytestmark = pytest.mark.gpu


def test_run_writes_u_and_iter(tmp_path):
    cfg = G.hetero3_small(1024)
    cfg["simulation"] = dict(cfg["simulation"], renderingsNum=2)
    m = _mat(cfg)
    m["scattering"]["type"] = _cell(cfg["scattering"]["type"])
    m["scattering"]["amplitudeFunction"] = _cell(list(cfg["scattering"]["amplitudeFunction"]))
    src = os.path.join(str(tmp_path), "cfg.mat")
    out = os.path.join(str(tmp_path), "r.mat")
    sio.savemat(src, {"config": m})
    assert M.main([src, "--out", out]) == 0
    r = sio.loadmat(out)
    u, it = sstmc3D(cfg)
    assert r["u"].shape == u.shape == (4, 4, 1, 2)
    assert np.array_equal(r["u"], u)
    assert np.array_equal(r["iter"].ravel(), np.ravel(it))
