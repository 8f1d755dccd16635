"""Helpers to drive the command-line host (temporalspecklescode_amd/sstmc_cli) from tests."""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "temporalspecklescode_amd", "sstmc_cli")


def to_json(obj):
    """Config dict (numpy arrays allowed) -> JSON-ready structure; n-d arrays become nested lists
    with the first index outermost, which the CLI reads back with the first index fastest."""
    if isinstance(obj, dict):
        return {k: to_json(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [to_json(v) for v in obj]
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    if isinstance(obj, np.generic):
        return obj.item()
    return obj


def run(cfg, tmpdir, *args, dims=None):
    path = os.path.join(str(tmpdir), "cfg.json")
    c = to_json(cfg)
    if dims is not None:
        c["dims"] = dims
    with open(path, "w") as f:
        json.dump(c, f)
    return subprocess.run([CLI, *args, path], capture_output=True, text=True, timeout=600)


def ensure_built():
    if not os.path.exists(CLI):
        from temporalspecklescode_amd import build
        build.build_cli()
