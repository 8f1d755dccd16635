"""The command-line host renders the same field as the Python host mirror (same C ABI)."""
import json
import os

import numpy as np
import pytest

from temporalspecklescode_amd import render
from temporalspecklescode_amd import workloads as W

import cli_util as U


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1", "c2", "confocal"])
def test_cli_render_matches_python(tmp_path, case):
    U.ensure_built()
    if case == "c1":
        cfg, dims = W.c1(paths=4096, batch=1024), 2
        cfg["simulation"]["renderingsNum"] = 2
    elif case == "c2":
        cfg, dims = W.c2(paths=2048, batch=1024, n=16), 3
    else:   # GaussianBeam illumination and view
        cfg, dims = W.confocal(focal_x=[-2.0, 0.0, 2.0], focal_z=[4.0, 12.0], tau_x=(0.0, 1.0), batch=512,
                               full_iteration=2), 3
    out = os.path.join(str(tmp_path), "u.npy")
    r = U.run(cfg, tmp_path, "--out", out, dims=dims)
    assert r.returncode == 0, r.stderr
    meta = json.loads(r.stdout)
    u_cli = np.load(out)
    u_py, it_py = render(cfg, dims)
    assert u_cli.shape == u_py.shape == tuple(meta["shape"])
    assert np.array_equal(u_cli, u_py)
    assert np.array_equal(np.asarray(meta["iter"]), it_py)
